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

#include <algorithm>

#include <array>
#include <map>
#include <mutex>

using namespace codens;

// Increment buffer view: element (e, n) at p[e*ld + n - off].  T = double
// for the FP64 phase, float for the FP32 defect-correction form.
template <typename T>
struct TView {
  T *p;
  long ld, off;
  __device__ inline T &at(int e, long n) const { return p[e * ld + n - off]; }
};
using DView = TView<double>;

// sn(m, d, c) of a block's metric slice in either precision
template <typename T>
__device__ inline T snt(const T *__restrict__ sn, long nnn, long m, int d, int c) {
  return sn[((long)c * 3 + d) * nnn + m];
}

template <typename T>
struct DFluxT {  // mass-flux increment (d of rho u, rho v, rho w) and dF
  T mx, my, mz, xx, xy, xz, yy, yz, zz, ex, ey, ez;
};
using DFlux = DFluxT<double>;

// implimf_3d.f90:434-465 for one cell: dF = F(v + d) - F(v), p(v) from the EOS
// as zpres.f90:72-73 computes it.  The velocities are formed with one
// reciprocal of each density instead of the reference's divisions (about
// 1e-16 relative per value; the sub-iteration result moves by ~3e-14).
template <typename T>
__device__ inline DFluxT<T> dflux(const Phys &ph, const T V[5], const T D[5]) {
  const T gam1 = (T)ph.gam1, pinfl = (T)ph.pinfl;
  T wi1 = V[0] + D[0], wi2 = V[1] + D[1], wi3 = V[2] + D[2], wi4 = V[3] + D[3], wi5 = V[4] + D[4];
  T ri = T(1) / wi1, r0 = T(1) / V[0];
  T ui = wi2 * ri, vi = wi3 * ri, wi = wi4 * ri;
  T pres = gam1 * (wi5 - T(.5) * wi1 * (ui * ui + vi * vi + wi * wi) - pinfl);
  T a = V[1] * r0, bb = V[2] * r0, c = V[3] * r0;
  T rhoe = V[4] - T(0.5) * (V[1] * a + V[2] * bb + V[3] * c);
  T p = gam1 * (rhoe - pinfl);
  T fxx = V[1] * a + p, fxy = V[2] * a, fxz = V[3] * a;
  T fyy = V[2] * bb + p, fyz = V[3] * bb, fzz = V[3] * c + p;
  T e5 = V[4] + p - pinfl;
  T fex = e5 * a, fey = e5 * bb, fez = e5 * c;
  DFluxT<T> r;
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

// The same increment dF = F(v + d) - F(v) in difference form, for the FP32
// sub-iterations: every term is a product with a change (dm - q drho)/(rho +
// drho) of a velocity, of the kinetic energy or of the pressure, so that its
// rounding is relative to |d|.  The form above subtracts two fluxes of size
// |F|, whose rounding in FP32 (6e-8 |F|) does not shrink with d and would
// hold the FP64 residual of the defect correction at about 1e-7 of its
// start.  Algebraically the same increment.
template <typename T>
__device__ inline DFluxT<T> dflux_delta(const Phys &ph, const T V[5], const T D[5]) {
  // no FMA contraction: the same rounding in every tile shape's kernel
#pragma clang fp contract(off)
  const T gam1 = (T)ph.gam1, pinfl = (T)ph.pinfl;
  T r0 = T(1) / V[0], rn = T(1) / (V[0] + D[0]);
  T u = V[1] * r0, v = V[2] * r0, w = V[3] * r0;
  T du = (D[1] - u * D[0]) * rn, dv = (D[2] - v * D[0]) * rn, dw = (D[3] - w * D[0]) * rn;
  T rho1 = V[0] + D[0];
  T q2 = u * u + v * v + w * w;
  T dq2 = T(2) * (u * du + v * dv + w * dw) + (du * du + dv * dv + dw * dw);
  T dk = T(0.5) * (D[0] * q2 + rho1 * dq2);                // change of rho |u|^2 / 2
  T p = gam1 * (V[4] - T(0.5) * V[0] * q2 - pinfl);
  T dp = gam1 * (D[4] - dk);
  // change of rho q_a q_b
  auto dprod = [&](T qa, T qb, T dqa, T dqb) {
#pragma clang fp contract(off)
    return D[0] * qa * qb + rho1 * (qa * dqb + dqa * qb + dqa * dqb);
  };
  T h0 = V[4] + p - pinfl, dh = D[4] + dp;                  // E + p - pinfl and its change
  DFluxT<T> r;
  r.mx = D[1]; r.my = D[2]; r.mz = D[3];
  r.xx = dprod(u, u, du, du) + dp;
  r.xy = dprod(u, v, du, dv);
  r.xz = dprod(u, w, du, dw);
  r.yy = dprod(v, v, dv, dv) + dp;
  r.yz = dprod(v, w, dv, dw);
  r.zz = dprod(w, w, dw, dw) + dp;
  r.ex = dh * (u + du) + h0 * du;
  r.ey = dh * (v + dv) + h0 * dv;
  r.ez = dh * (w + dw) + h0 * dw;
  return r;
}

// dF of the sub-iterations: the reference's form in FP64, the difference
// form in FP32
template <typename T>
__device__ inline DFluxT<T> dflux_sweep(const Phys &ph, const T V[5], const T D[5]) {
  if constexpr (sizeof(T) == sizeof(float)) return dflux_delta<T>(ph, V, D);
  else return dflux<T>(ph, V, D);
}

// implimf_eu.f90:289-312 for one cell: dF from the momenta with
// wi1 = 1/(rho + d1), the base flux with the caller's pressure p, and the z
// energy increment without its w factor, as the reference has it.
__device__ inline DFlux dflux_eu(const Phys &ph, const double V[5], const double D[5], double p) {
  const double gam1 = ph.gam1, pinfl = ph.pinfl;
  double wi1 = 1. / (V[0] + D[0]);
  double ui = V[1] + D[1], vi = V[2] + D[2], wi = V[3] + D[3], wi5 = V[4] + D[4];
  double pres = gam1 * (wi5 - .5 * wi1 * (ui * ui + vi * vi + wi * wi) - pinfl);
  double norm = 1. / V[0];
  double a = V[1] * norm, bb = V[2] * norm, c = V[3] * norm;
  double fxx = V[1] * a + p, fxy = V[2] * a, fxz = V[3] * a;
  double fyy = V[2] * bb + p, fyz = V[3] * bb, fzz = V[3] * c + p;
  double e5 = V[4] + p - pinfl;
  double fex = e5 * V[1] * norm, fey = e5 * V[2] * norm, fez = e5 * V[3] * norm;
  double h = wi5 + pres - pinfl;
  DFlux r;
  r.mx = D[1]; r.my = D[2]; r.mz = D[3];
  r.xx = wi1 * (ui * ui) + pres - fxx;
  r.xy = wi1 * ui * vi - fxy;
  r.xz = wi1 * ui * wi - fxz;
  r.ex = wi1 * ui * h - fex;
  r.yy = wi1 * (vi * vi) + pres - fyy;
  r.yz = wi1 * vi * wi - fyz;
  r.ey = wi1 * vi * h - fey;
  r.zz = wi1 * (wi * wi) + pres - fzz;
  r.ez = wi1 * h - fez;
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

#define CELL3(b) CELL3_AT(b, xcd_block())
// the same, for the tile XB of a block's own grid
#define CELL3_AT(b, XB)                              \
  int c[3];                                          \
  const dim3 xb = (XB);                              \
  c[0] = (b).i1 + xb.x * TX + threadIdx.x;     \
  c[1] = (b).j1 + xb.y * TY + threadIdx.y;     \
  c[2] = (b).k1 + xb.z;                        \
  if (c[0] >= (b).i2 || c[1] >= (b).j2) return;

__device__ inline double rv_of(const Phys &ph, const double *mu, const double *mut, double rinv, long n) {
  return ph.gam * (mu[n] / ph.pr + mut[n] / ph.prt) * rinv;
}

// implimf_3d.f90:113-224: spectral-radius coefficients of the faces and the
// diagonal; coefe[d] holds the low face of every real cell.  A workgroup
// owns a TX x TY tile of real cells and marches a slab of up to CK_KS
// k-planes (fewer on small blocks, slab_depth), so that each face is
// evaluated once.  Per plane, the velocity, sound speed, rv and
// volume of every cell of the tile and of its 1-cell ring (i0-1..i0+TX,
// j0-1..j0+TY) go to LDS; each thread evaluates the low i and j faces of its
// cell, the ring's threads the tile's high faces, and the k face below its
// cell from its own column's previous plane.  A face's coefficient is one
// expression of its two cells whichever tile evaluates it:
// 0.5 (|V.S| + |S| c) + (rv_lo + rv_hi) |S|^2 / (vol_lo + vol_hi), with V
// and c averaged (high + low) / 2 and no contraction.  Velocities and rv use
// one reciprocal of each density (the reference divides each time; the
// coefficients move by ~1e-16 relative).  A cell's diagonal adds its six
// faces in the order of implimf_3d.f90:206-215 (i low, i high, j low, j high,
// k low, k high) and is complete once the next plane has evaluated the k face
// above it.  Every global load of a plane is issued before its first barrier.
constexpr int CK_KS = 8;
constexpr int CK_PL = (TX + 2) * (TY + 2);

struct CoefCell {  // the per-cell operands of a face coefficient
  double u, v, w, c, rv, vol;
};

__device__ __forceinline__ CoefCell coef_cell(const Blk &b, const Phys &ph, const double *__restrict__ v, long ldv,
                                              const double *__restrict__ mu, const double *__restrict__ mut,
                                              const double *__restrict__ vol, const double *__restrict__ cson,
                                              int i, int j, int k) {
  long n = b.idx(i, j, k);
  double r = 1.0 / v[n];
  CoefCell q;
  q.u = v[ldv + n] * r;
  q.v = v[2 * ldv + n] * r;
  q.w = v[3 * ldv + n] * r;
  q.c = cson[n];
  q.rv = b.real(i, j, k) ? rv_of(ph, mu, mut, r, n) : 0.;
  q.vol = vol[n];
  return q;
}

// coefficient of the face between cells lo and hi (normal s of the face)
__device__ __forceinline__ double coef_face(const CoefCell &hi, const CoefCell &lo, double s1, double s2, double s3) {
#pragma clang fp contract(off)
  double cnds = s1 * s1 + s2 * s2 + s3 * s3;
  double uu = 0.5 * (hi.u + lo.u), vv = 0.5 * (hi.v + lo.v), ww = 0.5 * (hi.w + lo.w);
  double cc = 0.5 * (hi.c + lo.c), rsum = hi.rv + lo.rv, vsum = hi.vol + lo.vol;
  double vnf = uu * s1 + vv * s2 + ww * s3;
  return 0.5 * (fabs(vnf) + sqrt(cnds) * cc) + rsum * cnds / vsum;
}

// LDS of one workgroup: the plane's cell operands and the i and j face
// coefficients of the tile
struct CoefLds {
  double L[6][CK_PL];
  double CI[(TX + 1) * TY], CJ[TX * (TY + 1)];
};

// the tile xb (x, y: TX x TY columns, z: CK_KS-plane slab) of block b
__device__ __forceinline__ void coef_tiled_body(const Blk &b, dim3 xb, const Phys &ph, const double *__restrict__ v,
                                                long ldv, const double *__restrict__ mu,
                                                const double *__restrict__ mut, const double *__restrict__ dt,
                                                const double *__restrict__ vol, const double *__restrict__ cson,
                                                const double *__restrict__ sn, double dtpas, int kfmg,
                                                double *__restrict__ coefe, double *__restrict__ diag, int ks,
                                                CoefLds &S) {
#pragma clang fp contract(off)
  auto &L = S.L;
  double *CI = S.CI, *CJ = S.CJ;
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + TX * ty;
  const int i0 = b.i1 + xb.x * TX, j0 = b.j1 + xb.y * TY;
  const int kb = b.k1 + xb.z * ks, ke = min(kb + ks, b.k2);
  const int i = i0 + tx, j = j0 + ty;
  // own: a real cell; load: a real cell or the ghost past the block's high
  // i or j end, whose faces with the last real cells this thread evaluates
  // when the tile is partial (it is then inside the tile)
  const bool own = i < b.i2 && j < b.j2, load = i <= b.i2 && j <= b.j2;
  // ring cell of this thread (t < 2 TX + 2 TY): columns i0-1, i0+TX, rows
  // j0-1, j0+TY, loaded where it lies in the block's storage
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
  CoefCell below{}, cur{};
  if (own) below = coef_cell(b, ph, v, ldv, mu, mut, vol, cson, i, j, kb - 1);
  double part = 0.;  // the diagonal of the cell below, all but its k high face
  long mb = -1;
  // the tile's high face of a ring thread: i face at i0+TX of row j0+t
  // (t < TY), j face at j0+TY of column i0+t-TY (TY <= t < TY + TX), where it
  // bounds a real cell of the tile; its LDS slot and CI/CJ entry
  int hdir = -1, hs = 0, hc = 0, hfi = 0, hfj = 0;
  if (t < TY) {
    hfj = j0 + t; hfi = i0 + TX;
    if (hfi <= b.i2 && hfj < b.j2) { hdir = 0; hs = (t + 1) * (TX + 2) + TX + 1; hc = t * (TX + 1) + TX; }
  } else if (t < TY + TX) {
    hfi = i0 + t - TY; hfj = j0 + TY;
    if (hfj <= b.j2 && hfi < b.i2) { hdir = 1; hs = (TY + 1) * (TX + 2) + (t - TY) + 1; hc = TY * TX + (t - TY); }
  }
  // the low i face of this thread's cell where it bounds a real cell (rows
  // j < j2, up to the block's high boundary face i = i2), the low j face
  // likewise
  const bool fi = i <= b.i2 && j < b.j2, fj = i < b.i2 && j <= b.j2;
  for (int k = kb; k <= ke; k++) {
    const bool top = k == ke;  // the k face above the slab's last plane only
    // every global load of the plane, issued before its first barrier so
    // that their latencies overlap
    if (load) cur = coef_cell(b, ph, v, ldv, mu, mut, vol, cson, i, j, k);
    CoefCell rq{};
    const bool rl = !top && rneed;
    if (rl) rq = coef_cell(b, ph, v, ldv, mu, mut, vol, cson, ri, rj, k);
    long m = -1;
    double sk0 = 0., sk1 = 0., sk2 = 0., dtn = 1.;
    if (own) {
      m = b.idx(i, j, k) - b.npc;
      sk0 = snv(sn, b.nnn, m, 2, 0); sk1 = snv(sn, b.nnn, m, 2, 1); sk2 = snv(sn, b.nnn, m, 2, 2);
      if (!top) dtn = dt[m + b.npc];
    }
    const long mo = b.idx(min(i, b.i2), min(j, b.j2), k) - b.npc;
    double si0 = 0., si1 = 0., si2 = 0., sj0 = 0., sj1 = 0., sj2 = 0., sh0 = 0., sh1 = 0., sh2 = 0.;
    if (!top && fi) { si0 = snv(sn, b.nnn, mo, 0, 0); si1 = snv(sn, b.nnn, mo, 0, 1); si2 = snv(sn, b.nnn, mo, 0, 2); }
    if (!top && fj) { sj0 = snv(sn, b.nnn, mo, 1, 0); sj1 = snv(sn, b.nnn, mo, 1, 1); sj2 = snv(sn, b.nnn, mo, 1, 2); }
    if (!top && hdir >= 0) {
      long mf = b.idx(hfi, hfj, k) - b.npc;
      sh0 = snv(sn, b.nnn, mf, hdir, 0); sh1 = snv(sn, b.nnn, mf, hdir, 1); sh2 = snv(sn, b.nnn, mf, hdir, 2);
    }
    if (!top) {
      if (load) {
        L[0][slot] = cur.u; L[1][slot] = cur.v; L[2][slot] = cur.w;
        L[3][slot] = cur.c; L[4][slot] = cur.rv; L[5][slot] = cur.vol;
      }
      if (rl) {
        L[0][rslot] = rq.u; L[1][rslot] = rq.v; L[2][rslot] = rq.w;
        L[3][rslot] = rq.c; L[4][rslot] = rq.rv; L[5][rslot] = rq.vol;
      }
    }
    __syncthreads();
    double ckl = 0.;
    if (own) {
      ckl = coef_face(cur, below, sk0, sk1, sk2);
      if (mb >= 0) {  // the cell below is complete
        double dg = part + ckl;
        if (kfmg == 3) dg = dg + 1.5 * below.vol / dtpas;
        diag[mb] = dg;
      }
    }
    if (!top) {
      auto lds = [&](int s) {
        CoefCell q;
        q.u = L[0][s]; q.v = L[1][s]; q.w = L[2][s]; q.c = L[3][s]; q.rv = L[4][s]; q.vol = L[5][s];
        return q;
      };
      if (fi) CI[ty * (TX + 1) + tx] = coef_face(lds(slot), lds(slot - 1), si0, si1, si2);
      if (fj) CJ[ty * TX + tx] = coef_face(lds(slot), lds(slot - (TX + 2)), sj0, sj1, sj2);
      if (hdir == 0) CI[hc] = coef_face(lds(hs), lds(hs - 1), sh0, sh1, sh2);
      else if (hdir == 1) CJ[hc] = coef_face(lds(hs), lds(hs - (TX + 2)), sh0, sh1, sh2);
    }
    __syncthreads();
    if (!top && own) {
      double c0 = CI[ty * (TX + 1) + tx], c1 = CI[ty * (TX + 1) + tx + 1];
      double c2 = CJ[ty * TX + tx], c3 = CJ[(ty + 1) * TX + tx];
      coefe[m] = c0;
      coefe[b.nnn + m] = c2;
      coefe[2 * b.nnn + m] = ckl;
      part = cur.vol / dtn;
      part = part + c0 + c1 + c2 + c3 + ckl;
      mb = m;
    }
    below = cur;
  }
}

__global__ void __launch_bounds__(TX *TY) k_impl_coef_tiled(Blk b, Phys ph, const double *__restrict__ v, long ldv,
                                                            const double *__restrict__ mu,
                                                            const double *__restrict__ mut,
                                                            const double *__restrict__ dt,
                                                            const double *__restrict__ vol,
                                                            const double *__restrict__ cson,
                                                            const double *__restrict__ sn, double dtpas, int kfmg,
                                                            double *__restrict__ coefe, double *__restrict__ diag,
                                                            int ks) {
  __shared__ CoefLds S;
  coef_tiled_body(b, xcd_block(), ph, v, ldv, mu, mut, dt, vol, cson, sn, dtpas, kfmg, coefe, diag, ks, S);
}

// implimf_eu.f90:124-170: coefficients of the Euler implicit phase,
// coefe = 0.5 |V.S| + |S| c and diag = vol (fact + 1/dt) + sum of the face
// coefficients, direction by direction; fact = 1.5/dt1min when kfmg = 3.
__global__ void __launch_bounds__(TX *TY) k_impl_coef_eu(Blk b, const double *__restrict__ v, long ldv,
                                                         const double *__restrict__ dt,
                                                         const double *__restrict__ vol,
                                                         const double *__restrict__ cson,
                                                         const double *__restrict__ sn, double fact,
                                                         double *__restrict__ coefe, double *__restrict__ diag) {
  CELL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double norm = 1. / v[n];
  double u0 = v[ldv + n] * norm, v0 = v[2 * ldv + n] * norm, w0 = v[3 * ldv + n] * norm;
  double dg = vol[n] * (fact + 1. / dt[n]);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    long inc = b.inc(d);
    double cf[2];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      long nn = side == 0 ? n - inc : n + inc;
      long mf = side == 0 ? m : m + inc;
      double s1 = snv(sn, b.nnn, mf, d, 0), s2 = snv(sn, b.nnn, mf, d, 1), s3 = snv(sn, b.nnn, mf, d, 2);
      double cnds = s1 * s1 + s2 * s2 + s3 * s3;
      double norm1 = 1. / v[nn];
      double un = v[ldv + nn] * norm1, vn_ = v[2 * ldv + nn] * norm1, wn = v[3 * ldv + nn] * norm1;
      double uu, vv, ww, cc;
      if (side == 0) {  // face m: cells n (own) and n-inc
        uu = 0.5 * (u0 + un); vv = 0.5 * (v0 + vn_); ww = 0.5 * (w0 + wn); cc = 0.5 * (cson[n] + cson[nn]);
      } else {          // face m+inc: cells n+inc and n (own)
        uu = 0.5 * (un + u0); vv = 0.5 * (vn_ + v0); ww = 0.5 * (wn + w0); cc = 0.5 * (cson[nn] + cson[n]);
      }
      cf[side] = 0.5 * fabs(uu * s1 + vv * s2 + ww * s3) + sqrt(cnds) * cc;
    }
    coefe[d * b.nnn + m] = cf[0];
    dg = dg + cf[0] + cf[1];
  }
  diag[m] = dg;
}

// ---------------------------------------------------------------------------
// Tiled sub-iteration.  A workgroup owns an SW_TX x SW_TY column tile of real
// cells over SW_KS planes and marches k: the 14-vector q of a cell (mass-flux
// increments d2..d4, the nine dF, d1, d5) is formed once per sweep and kept
// in LDS for the i/j neighbours of its plane and in registers for the k
// neighbours; the k face term is carried from one plane to the next.
// LDS: 14 x (SW_TY+2) x (SW_TX+2) doubles = 44,352 bytes.  A 64 x 4 tile of
// 256 threads runs two workgroups per CU at its ~237 VGPRs (one workgroup
// waits at its plane barrier while the other computes): 125.6 us on the
// 2.1M-cell block against 133.1 for 64 x 8, despite its larger halo.
#ifndef CODENS_SW_TY
#define CODENS_SW_TY 4
#endif
#ifndef CODENS_SW_KS
#define CODENS_SW_KS 8
#endif
constexpr int SW_TX = 64, SW_TY = CODENS_SW_TY, SW_KS = CODENS_SW_KS;
static_assert(2 * SW_TX + 2 * SW_TY <= SW_TX * SW_TY, "one halo cell per thread at most");
constexpr int SW_HALO = 2 * SW_TX + 2 * SW_TY;

template <typename T>
struct Q14T {
  T q[14];
};
using Q14 = Q14T<double>;

template <typename T, class DIN>
__device__ inline void q_of(const Phys &ph, const T *__restrict__ v, long ldv, const DIN &d, long n, Q14T<T> &Q) {
  T V[5], D[5];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    V[e] = v[e * ldv + n];
    D[e] = d.at(e, n);
  }
  DFluxT<T> F = dflux_sweep<T>(ph, V, D);
  Q.q[0] = F.mx; Q.q[1] = F.my; Q.q[2] = F.mz;
  Q.q[3] = F.xx; Q.q[4] = F.xy; Q.q[5] = F.xz; Q.q[6] = F.yy; Q.q[7] = F.yz; Q.q[8] = F.zz;
  Q.q[9] = F.ex; Q.q[10] = F.ey; Q.q[11] = F.ez;
  Q.q[12] = D[0]; Q.q[13] = D[4];
}

// the same 14-vector for implimf_eu (its own flux increments, caller's p)
template <class DIN>
__device__ inline void q_of_eu(const Phys &ph, const double *__restrict__ v, long ldv, const DIN &d,
                               const double *__restrict__ ps, long n, Q14 &Q) {
  double V[5], D[5];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    V[e] = v[e * ldv + n];
    D[e] = d.at(e, n);
  }
  DFlux F = dflux_eu(ph, V, D, ps[n]);
  Q.q[0] = F.mx; Q.q[1] = F.my; Q.q[2] = F.mz;
  Q.q[3] = F.xx; Q.q[4] = F.xy; Q.q[5] = F.xz; Q.q[6] = F.yy; Q.q[7] = F.yz; Q.q[8] = F.zz;
  Q.q[9] = F.ex; Q.q[10] = F.ey; Q.q[11] = F.ez;
  Q.q[12] = D[0]; Q.q[13] = D[4];
}

template <bool EU, typename T, class DIN>
__device__ inline void q_model(const Phys &ph, const T *__restrict__ v, long ldv, const DIN &d,
                               const double *__restrict__ ps, long n, Q14T<T> &Q) {
  if constexpr (EU) q_of_eu(ph, v, ldv, d, ps, n, Q);
  else q_of<T, DIN>(ph, v, ldv, d, n, Q);
}

template <typename T>
__device__ inline void q_zero(Q14T<T> &Q) {
#pragma unroll
  for (int c = 0; c < 14; c++) Q.q[c] = T(0);
}

// The same 14-vector in two phases, so that the loads of several cells can
// be in flight together: the inputs of a cell (zero for a cell outside the
// real region), then q from them.
template <typename T>
struct VD {
  T V[5], D[5];
  double p;
};

template <bool EU, typename T, class DIN>
__device__ inline void vd_load(const T *__restrict__ v, long ldv, const DIN &d, const double *__restrict__ ps,
                               long n, bool real, VD<T> &x) {
#pragma unroll
  for (int e = 0; e < 5; e++) {
    x.V[e] = real ? v[e * ldv + n] : T(1);
    x.D[e] = real ? d.at(e, n) : T(0);
  }
  x.p = (EU && real) ? ps[n] : 0.;
}

template <bool EU, typename T>
__device__ inline void vd_q(const Phys &ph, const VD<T> &x, bool real, Q14T<T> &Q) {
  if (!real) {
    q_zero(Q);
    return;
  }
  DFluxT<T> F;
  if constexpr (EU) F = dflux_eu(ph, x.V, x.D, x.p);
  else F = dflux_sweep<T>(ph, x.V, x.D);
  Q.q[0] = F.mx; Q.q[1] = F.my; Q.q[2] = F.mz;
  Q.q[3] = F.xx; Q.q[4] = F.xy; Q.q[5] = F.xz; Q.q[6] = F.yy; Q.q[7] = F.yz; Q.q[8] = F.zz;
  Q.q[9] = F.ex; Q.q[10] = F.ey; Q.q[11] = F.ez;
  Q.q[12] = x.D[0]; Q.q[13] = x.D[4];
}


// d of component e from the 14-vector
template <typename T>
__device__ inline T q_d(const T *q, int e) {
  return e == 0 ? q[12] : (e == 4 ? q[13] : q[e - 1]);
}

// tn of a face between cells A and B (implimf_3d.f90:276-292)
template <typename T>
__device__ inline void q_tn(const T *A, const T *B, T s1, T s2, T s3, T tn[5]) {
  const T h = T(0.5);
  tn[0] = h * (A[0] + B[0]) * s1 + h * (A[1] + B[1]) * s2 + h * (A[2] + B[2]) * s3;
  tn[1] = h * (A[3] + B[3]) * s1 + h * (A[4] + B[4]) * s2 + h * (A[5] + B[5]) * s3;
  tn[2] = h * (A[4] + B[4]) * s1 + h * (A[6] + B[6]) * s2 + h * (A[7] + B[7]) * s3;
  tn[3] = h * (A[5] + B[5]) * s1 + h * (A[7] + B[7]) * s2 + h * (A[8] + B[8]) * s3;
  tn[4] = h * (A[9] + B[9]) * s1 + h * (A[10] + B[10]) * s2 + h * (A[11] + B[11]) * s3;
}

// First sub-iteration: d and dF are zero, so d = -(u (+ff)) / diag.
template <typename T>
__device__ __forceinline__ void sweep1_body(const Blk &b, dim3 tile, int ityprk, const T *__restrict__ u,
                                            const T *__restrict__ ff, long ldv, const T *__restrict__ diag,
                                            const TView<T> &dnew) {
  CELL3_AT(b, tile)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  T dg = diag[m];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    T r = ityprk == 0 ? -u[e * ldv + n] : -u[e * ldv + n] - ff[e * ldv + n];
    dnew.at(e, n) = r / dg;
  }
}

template <typename T>
__global__ void __launch_bounds__(TX *TY) k_impl_sweep1(Blk b, int ityprk, const T *__restrict__ u,
                                                        const T *__restrict__ ff, long ldv,
                                                        const T *__restrict__ diag, TView<T> dnew) {
  sweep1_body<T>(b, xcd_block(), ityprk, u, ff, ldv, diag, dnew);
}

constexpr int SW_PL = (SW_TY + 2) * (SW_TX + 2);  // LDS plane stride between the 14 components

// face term of the face between the LDS cells a (own side) and o (other side),
// components read straight from LDS: tn = 0.5 (q_a + q_o) . S
template <int P, typename T>
__device__ inline T lds_tn(const T *a, const T *o, int e, T s1, T s2, T s3) {
  const T h = T(0.5);
  switch (e) {
    case 0: return h * (a[0] + o[0]) * s1 + h * (a[P] + o[P]) * s2 + h * (a[2 * P] + o[2 * P]) * s3;
    case 1: return h * (a[3 * P] + o[3 * P]) * s1 + h * (a[4 * P] + o[4 * P]) * s2 + h * (a[5 * P] + o[5 * P]) * s3;
    case 2: return h * (a[4 * P] + o[4 * P]) * s1 + h * (a[6 * P] + o[6 * P]) * s2 + h * (a[7 * P] + o[7 * P]) * s3;
    case 3: return h * (a[5 * P] + o[5 * P]) * s1 + h * (a[7 * P] + o[7 * P]) * s2 + h * (a[8 * P] + o[8 * P]) * s3;
    default: return h * (a[9 * P] + o[9 * P]) * s1 + h * (a[10 * P] + o[10 * P]) * s2 + h * (a[11 * P] + o[11 * P]) * s3;
  }
}
template <int P, typename T>
__device__ inline T lds_d(const T *a, int e) {
  return e == 0 ? a[12 * P] : (e == 4 ? a[13 * P] : a[(e - 1) * P]);
}

// Workgroup barrier that orders only the LDS traffic of the plane exchange.
__device__ inline void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#ifdef CODENS_SW_WAVES
#define SW_ATTR __attribute__((amdgpu_waves_per_eu(CODENS_SW_WAVES, CODENS_SW_WAVES)))
#else
#define SW_ATTR
#endif
// EU selects the flux increments of implimf_eu (which read ps); the face
// terms, coefficients and update are common to both implicit phases.
template <bool EU, typename T, int TXS, int TYS, class DIN>
__device__ __forceinline__ void sweep_body(const Blk b, int lg, int ntx, int nty, int ks, const Phys ph, int ityprk,
                                           const T *__restrict__ u, const T *__restrict__ ff,
                                           const T *__restrict__ v, long ldv, const double *__restrict__ ps,
                                           const T *__restrict__ sn, const T *__restrict__ coefe,
                                           const T *__restrict__ diag, const DIN dold, const TView<T> dnew,
                                           T *sq) {
  constexpr int PL = (TYS + 2) * (TXS + 2), HALO = 2 * TXS + 2 * TYS;
  static_assert(HALO <= TXS * TYS, "one halo cell per thread at most");
#include "sweep_tile.inc"
}

template <bool EU, typename T, int TXS = SW_TX, int TYS = SW_TY>
__global__ void __launch_bounds__(TXS *TYS) SW_ATTR k_impl_sweep_tiled(Blk b, Phys ph, int ityprk, int ntx, int nty,
                                                                   int nkz, int ks, const T *__restrict__ u,
                                                                   const T *__restrict__ ff,
                                                                   const T *__restrict__ v, long ldv,
                                                                   const double *__restrict__ ps,
                                                                   const T *__restrict__ sn,
                                                                   const T *__restrict__ coefe,
                                                                   const T *__restrict__ diag, TView<T> dold,
                                                                   TView<T> dnew, int rev) {
  // XCD-aware tile order: the hardware deals workgroups to the 8 XCDs round
  // robin; give each XCD a contiguous run of tiles so that j/k neighbours
  // share its L2.  rev runs each XCD's tiles in reverse: alternate
  // sub-iterations then begin where the previous one ended, on the data it
  // left in the last-level cache (the tile order does not change any value).
  constexpr int PL = (TYS + 2) * (TXS + 2), HALO = 2 * TXS + 2 * TYS;
  static_assert(HALO <= TXS * TYS, "one halo cell per thread at most");
  __shared__ T sq[14 * PL];
  int nwg = ntx * nty * nkz;
  int hw = blockIdx.x;
  int lg = (nwg % 8 == 0) ? (hw % 8) * (nwg / 8) + hw / 8 : hw;
  if (rev) lg = nwg - 1 - lg;
#include "sweep_tile.inc"
}

// ---------------------------------------------------------------------------
// Tile shape of the tiled sub-iteration, per block shape: 64 x 4, 32 x 8 or
// 16 x 16, all of 256 threads.  Every cell's arithmetic is the same in all,
// so the increments are identical bit for bit; what differs is how many
// lanes idle on the i and j edges of a block, the LDS halo per cell and the
// length of the rows each wave reads.  By default the first call for a
// block shape times each on the caller's own data (one untimed and two
// timed launches each, on the caller's stream, writing the same result the
// real launch writes next) and keeps the fastest; while the stream is being
// captured into a graph, an unseen block shape takes tile_model's choice.
// codens_sweep_tile forces one shape.
namespace {
// shapes 0 (SW_TX x SW_TY), 1 (32 x 2 SW_TY) and 2 (16 x 16), all of 256
// threads; the all-blocks launch holds the first two (N_BATCH_TILES)
constexpr int N_SWEEP_TILES = 3, N_BATCH_TILES = 2;
std::mutex g_tile_mu;
int g_tile_mode = -1;
std::map<std::array<int, 5>, int> g_tile_best;

inline int tile_tx(int tile) { return tile == 2 ? 16 : (tile == 1 ? 32 : SW_TX); }
inline int tile_ty(int tile) { return tile == 2 ? 16 : (tile == 1 ? 2 * SW_TY : SW_TY); }

// The shape with the least modelled time among the first n: the lanes the
// shape launches over the block's i-j extent (idle lanes past its edges
// included) times the shape's measured time per lane on a block that none
// of them leaves idle lanes in (256 x 128 x 64 on MI355X: 64 x 4 134.7 us,
// 32 x 8 139.2 us, 16 x 16 154.3 us).  The narrower shapes read more halo
// cells per cell and shorter runs of each row.
int tile_model(const Blk &b, int n = N_BATCH_TILES) {
  static const double per_lane[N_SWEEP_TILES] = {1.0, 139.2 / 134.7, 154.3 / 134.7};
  int nx = b.i2 - b.i1, ny = b.j2 - b.j1;
  double best = 0.;
  int pick = 0;
  for (int t = 0; t < n; t++) {
    int tx = tile_tx(t), ty = tile_ty(t);
    double lanes = (double)((nx + tx - 1) / tx * tx) * ((ny + ty - 1) / ty * ty);
    double cost = lanes * per_lane[t];
    if (t == 0 || cost < best) { best = cost; pick = t; }
  }
  return pick;
}

template <bool EU, typename T>
void launch_sweep(int tile, const Blk &b, hipStream_t st, Phys ph, int ityprk, const T *u, const T *ff, const T *v,
                  long ldv, const double *ps, const T *sn, const T *coefe, const T *diag, TView<T> in,
                  TView<T> out, int rev = 0) {
  const int tx = tile_tx(tile), ty = tile_ty(tile);
  const int ntx0 = (b.i2 - b.i1 + tx - 1) / tx, nty0 = (b.j2 - b.j1 + ty - 1) / ty;
  const int ks = slab_depth((long)ntx0 * nty0, b.k2 - b.k1, SW_KS);
  const int nkz = (b.k2 - b.k1 + ks - 1) / ks;
  if (tile == 2) {
    constexpr int X = 16, Y = 16;
    int ntx = (b.i2 - b.i1 + X - 1) / X, nty = (b.j2 - b.j1 + Y - 1) / Y;
    hipLaunchKernelGGL((k_impl_sweep_tiled<EU, T, X, Y>), dim3(ntx * nty * nkz), dim3(X, Y), 0, st, b, ph,
                       ityprk, ntx, nty, nkz, ks, u, ff, v, ldv, ps, sn, coefe, diag, in, out, rev);
  } else if (tile == 1) {
    constexpr int X = 32, Y = 2 * SW_TY;
    int ntx = (b.i2 - b.i1 + X - 1) / X, nty = (b.j2 - b.j1 + Y - 1) / Y;
    hipLaunchKernelGGL((k_impl_sweep_tiled<EU, T, X, Y>), dim3(ntx * nty * nkz), dim3(X, Y), 0, st, b, ph,
                       ityprk, ntx, nty, nkz, ks, u, ff, v, ldv, ps, sn, coefe, diag, in, out, rev);
  } else {
    int ntx = (b.i2 - b.i1 + SW_TX - 1) / SW_TX, nty = (b.j2 - b.j1 + SW_TY - 1) / SW_TY;
    hipLaunchKernelGGL((k_impl_sweep_tiled<EU, T, SW_TX, SW_TY>), dim3(ntx * nty * nkz), dim3(SW_TX, SW_TY),
                       0, st, b, ph, ityprk, ntx, nty, nkz, ks, u, ff, v, ldv, ps, sn, coefe, diag, in, out, rev);
  }
}

template <bool EU, typename T>
int pick_tile(const Blk &b, hipStream_t st, Phys ph, int ityprk, const T *u, const T *ff, const T *v, long ldv,
              const double *ps, const T *sn, const T *coefe, const T *diag, TView<T> in, TView<T> out) {
  std::array<int, 5> key = {b.i2 - b.i1, b.j2 - b.j1, b.k2 - b.k1, (int)EU, (int)sizeof(T)};
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    if (g_tile_mode >= 0) return g_tile_mode;
    auto it = g_tile_best.find(key);
    if (it != g_tile_best.end()) return it->second;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return tile_model(b);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return tile_model(b);
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return tile_model(b);
  }
  int pick = tile_model(b);
  float best = -1.f;
  for (int t = 0; t < N_SWEEP_TILES; t++) {
    launch_sweep<EU, T>(t, b, st, ph, ityprk, u, ff, v, ldv, ps, sn, coefe, diag, in, out);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < 2; r++)
      launch_sweep<EU, T>(t, b, st, ph, ityprk, u, ff, v, ldv, ps, sn, coefe, diag, in, out);
    (void)hipEventRecord(e1, st);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      pick = tile_model(b);
      best = -1.f;
      break;
    }
    if (best < 0.f || ms < best) { best = ms; pick = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tile_best[key] = pick;
  return pick;
}
}  // namespace

// tile of a block without timing anything: the forced mode, the choice
// already timed for its shape, or the model's
static int known_tile(const Blk &b, bool eu, int tsize) {
  std::array<int, 5> key = {b.i2 - b.i1, b.j2 - b.j1, b.k2 - b.k1, eu ? 1 : 0, tsize};
  std::lock_guard<std::mutex> lk(g_tile_mu);
  if (g_tile_mode >= 0) return g_tile_mode;
  auto it = g_tile_best.find(key);
  return it != g_tile_best.end() ? it->second : tile_model(b);
}

extern "C" int codens_sweep_tile(int32_t mode, int32_t *previous) {
  if (mode < -1 || mode >= N_SWEEP_TILES) {
    codens_set_error("codens_sweep_tile: mode %d is not -1 (timed choice), 0 (64 x 4), 1 (32 x 8) or 2 (16 x 16)",
                     mode);
    return CODENS_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_tile_mu);
  if (previous) *previous = g_tile_mode;
  g_tile_mode = mode;
  return CODENS_OK;
}

extern "C" int codens_sweep_tile_of(const codens_block *cb, int32_t eu, int32_t fp32, int32_t *tile) {
  int rc = codens_check_block(cb, "codens_sweep_tile_of");
  if (rc) return rc;
  if (!tile) {
    codens_set_error("codens_sweep_tile_of: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  std::array<int, 5> key = {b.i2 - b.i1, b.j2 - b.j1, b.k2 - b.k1, eu ? 1 : 0, fp32 ? 4 : 8};
  std::lock_guard<std::mutex> lk(g_tile_mu);
  auto it = g_tile_best.find(key);
  *tile = g_tile_mode >= 0 ? g_tile_mode : (it != g_tile_best.end() ? it->second : -1);
  return CODENS_OK;
}

// implimf_3d.f90:473-486
__device__ __forceinline__ void update_body(const Blk &b, dim3 tile, double *v, long ldv,
                                            const double *__restrict__ d, int nvar) {
  CELL3_AT(b, tile)
  long n = b.idx(c[0], c[1], c[2]);
  for (int e = 0; e < nvar; e++) v[e * ldv + n] = v[e * ldv + n] + d[e * ldv + n];
}

__global__ void __launch_bounds__(TX *TY) k_impl_update(Blk b, double *v, long ldv, const double *__restrict__ d,
                                                        int nvar) {
  update_body(b, xcd_block(), v, ldv, d, nvar);
}

__global__ void k_zero5(double *d, long ldv, long n0, long n1) {
  long n = n0 + blockIdx.x * 256L + threadIdx.x;
  if (n > n1) return;
#pragma unroll
  for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
}

// Shared host sequence of both implicit phases: zero d, coefficients, lmx
// sub-iterations (ping-pong between the caller's d and the workspace, the
// last one landing in d), then v += d on nvar planes.
// ws_ext: scratch of 9 nnn doubles supplied by the caller (then nothing is
// allocated, no timing event is created and the sub-iteration tile is the
// forced or already chosen one, else the model's); nullptr: the library's
// workspace.
static int implimf_common(bool eu, const codens_block *cb, const codens_phys *cp, const double *u, const double *dt,
                          double *v, int64_t ldv, int nvar, double *d, const double *ff, const double *mu,
                          const double *mut, int32_t lmx, const double *sn, const double *vol, double dtpas,
                          double dt1min, int32_t kfmg, int32_t ityprk, const double *ps, const double *cson,
                          void *stream, const char *what, double *ws_ext = nullptr) {
  int rc = codens_check_block(cb, what);
  if (rc) return rc;
  if (!cp || !u || !dt || !v || !d || (!eu && (!mu || !mut)) || (eu && !ps) || !sn || !vol || !cson ||
      (ityprk && !ff) || lmx < 0 || nvar < 5) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  if (eu && kfmg == 3 && !(dt1min > 0.)) {
    codens_set_error("%s: dt1min must be positive when kfmg = 3", what);
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Phys ph = make_phys(cp);
  hipStream_t st = (hipStream_t)stream;
  const bool fixed = ws_ext != nullptr;
  double *ws = fixed ? ws_ext : (double *)codens_workspace(sizeof(double) * 9 * b.nnn, 2);
  if (!ws) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  double *coefe = ws, *diag = ws + 3 * b.nnn, *W = ws + 4 * b.nnn;
  long n0 = b.npc, n1 = b.npc + b.nnn - 1;
  auto tname = [fixed](const char *n) { return fixed ? (const char *)nullptr : n; };
  // d = 0 on the ghosts; the last sub-iteration writes every real cell
  if (lmx >= 1) codens_zero_shell(b, n0, n1, d, (long)ldv, 5, st);
  else hipLaunchKernelGGL(k_zero5, dim3((b.nnn + 255) / 256), dim3(256), 0, st, d, (long)ldv, n0, n1);
  dim3 g = cell_grid(b), t(TX, TY);
  {
    CodensTimed _tm(tname("implimf_coef"), st);
    if (eu) {
      hipLaunchKernelGGL(k_impl_coef_eu, g, t, 0, st, b, v, (long)ldv, dt, vol, cson, sn,
                         kfmg == 3 ? 1.5 / dt1min : 0., coefe, diag);
    } else {
      const int ks = slab_depth((long)g.x * g.y, b.k2 - b.k1, CK_KS);
      dim3 gc(g.x, g.y, (unsigned)((b.k2 - b.k1 + ks - 1) / ks));
      hipLaunchKernelGGL(k_impl_coef_tiled, gc, t, 0, st, b, ph, v, (long)ldv, mu, mut, dt, vol, cson, sn, dtpas,
                         (int)kfmg, coefe, diag, ks);
    }
  }
  DView A = {d, (long)ldv, 0}, B = {W, b.nnn, b.npc};
  int tile = 0;
  // sub-iteration 1 starts from d = 0: d = -u/diag (k_impl_sweep1); the
  // others are tiled sweeps, alternate ones in reverse tile order
  for (int s = 1; s <= lmx; s++) {
    DView out = ((lmx - s) % 2 == 0) ? A : B;
    DView in = ((lmx - s) % 2 == 0) ? B : A;
    const double *fx = ityprk ? ff : u;
    if (s == 1) {
      CodensTimed _tm(tname("implimf_sweep1"), st);
      hipLaunchKernelGGL(k_impl_sweep1<double>, g, t, 0, st, b, (int)ityprk, u, fx, (long)ldv, diag, out);
    } else {
      const double *psx = eu ? ps : v;
      if (s == 2 && fixed)
        tile = known_tile(b, eu, (int)sizeof(double));
      else if (s == 2)
        tile = eu ? pick_tile<true, double>(b, st, ph, (int)ityprk, u, fx, v, (long)ldv, psx, sn, coefe, diag, in, out)
                  : pick_tile<false, double>(b, st, ph, (int)ityprk, u, fx, v, (long)ldv, psx, sn, coefe, diag, in,
                                             out);
      CodensTimed _tm(tname("implimf_sweep"), st);
      if (eu)
        launch_sweep<true, double>(tile, b, st, ph, (int)ityprk, u, fx, v, (long)ldv, psx, sn, coefe, diag, in, out,
                                   s % 2 == 0);
      else
        launch_sweep<false, double>(tile, b, st, ph, (int)ityprk, u, fx, v, (long)ldv, psx, sn, coefe, diag, in, out,
                                    s % 2 == 0);
    }
  }
  {
    CodensTimed _tm(tname("implimf_update"), st);
    hipLaunchKernelGGL(k_impl_update, g, t, 0, st, b, v, (long)ldv, d, nvar);
  }
  return codens_check_launch(what);
}

int codens_implimf_3d_with(const codens_block *b, const codens_phys *ph, const double *u, const double *dt, double *v,
                           int64_t ldv, double *d, const double *ff, const double *mu, const double *mut, int32_t lmx,
                           const double *sn_blk, const double *vol, double dtpas, int32_t kfmg, int32_t ityprk,
                           const double *cson, void *stream, double *ws, const char *what) {
  return implimf_common(false, b, ph, u, dt, v, ldv, 5, d, ff, mu, mut, lmx, sn_blk, vol, dtpas, 1.0, kfmg, ityprk,
                        nullptr, cson, stream, what, ws);
}

// ---------------------------------------------------------------------------
// implimf_3d for all blocks of a layout.  Each pass (zero d, coefficients,
// first sub-iteration, the lmx-1 tiled sub-iterations, update) is one
// flattened launch over every block; a device table gives each block its
// extents and its first workgroup per pass, and a workgroup runs the
// per-block kernel's own body on its block's tile.  Block l's coefficients,
// diagonal and ping-pong increment (9 planes of nnn) start at 9 npc of the
// workspace.  The sub-iterations of small blocks then share the chip instead
// of each filling a few CUs.
enum { IB_ZERO, IB_CELL, IB_SWEEP, IB_COEF, IB_N };

struct ImplEnt {
  Blk b;
  unsigned start[IB_N];         // first workgroup of this block, per pass
  unsigned gx[IB_N], gy[IB_N];  // this block's grid, per pass
  int tile;                     // sub-iteration tile shape of this block (0: 64 x 4, 1: 32 x 8)
  int ks_coef, ks_sweep;        // k planes per workgroup
};

__global__ void __launch_bounds__(256) k_zero5_batch(const ImplEnt *__restrict__ tab, int nb, double *d, long ldv) {
  const ImplEnt &E = tab[tab_find(tab, nb, IB_ZERO, blockIdx.x)];
  const long n0 = E.b.npc, n = n0 + (blockIdx.x - E.start[IB_ZERO]) * 256L + threadIdx.x;
  if (n >= n0 + E.b.nnn) return;
#pragma unroll
  for (int e = 0; e < 5; e++) d[e * ldv + n] = 0.;
}

__global__ void __launch_bounds__(TX *TY) k_impl_coef_batch(const ImplEnt *__restrict__ tab, int nb, Phys ph,
                                                            const double *__restrict__ v, long ldv,
                                                            const double *__restrict__ mu,
                                                            const double *__restrict__ mut,
                                                            const double *__restrict__ dt,
                                                            const double *__restrict__ vol,
                                                            const double *__restrict__ cson,
                                                            const double *__restrict__ sn_all, double dtpas, int kfmg,
                                                            double *ws) {
  __shared__ CoefLds S;
  const unsigned wg = xcd_linear();
  const ImplEnt &E = tab[tab_find(tab, nb, IB_COEF, wg)];
  const Blk b = E.b;
  double *c = ws + 9 * b.npc;
  coef_tiled_body(b, tab_tile(E, IB_COEF, wg), ph, v, ldv, mu, mut, dt, vol, cson, sn_all + 9 * b.npc, dtpas, kfmg,
                  c, c + 3 * b.nnn, E.ks_coef, S);
}

__global__ void __launch_bounds__(TX *TY) k_impl_sweep1_batch(const ImplEnt *__restrict__ tab, int nb, int ityprk,
                                                              const double *__restrict__ u,
                                                              const double *__restrict__ ff, long ldv, double *ws,
                                                              double *d, int out_d) {
  const unsigned wg = xcd_linear();
  const ImplEnt &E = tab[tab_find(tab, nb, IB_CELL, wg)];
  const Blk b = E.b;
  double *c = ws + 9 * b.npc;
  DView out = out_d ? DView{d, ldv, 0} : DView{c + 4 * b.nnn, b.nnn, b.npc};
  sweep1_body<double>(b, tab_tile(E, IB_CELL, wg), ityprk, u, ff, ldv, c + 3 * b.nnn, out);
}

__global__ void __launch_bounds__(SW_TX *SW_TY) k_impl_sweep_batch(const ImplEnt *__restrict__ tab, int nb, Phys ph,
                                                                   int ityprk, const double *__restrict__ u,
                                                                   const double *__restrict__ ff,
                                                                   const double *__restrict__ v, long ldv,
                                                                   const double *__restrict__ sn_all, double *ws,
                                                                   double *d, int out_d, int rev) {
  constexpr int PL0 = (SW_TY + 2) * (SW_TX + 2), PL1 = (2 * SW_TY + 2) * (32 + 2);
  __shared__ double sq[14 * (PL0 > PL1 ? PL0 : PL1)];  // one buffer for both shapes
  const unsigned wg = rev ? gridDim.x - 1 - xcd_linear() : xcd_linear();
  const ImplEnt &E = tab[tab_find(tab, nb, IB_SWEEP, wg)];
  const Blk b = E.b;
  double *c = ws + 9 * b.npc;
  DView A = {d, ldv, 0}, B = {c + 4 * b.nnn, b.nnn, b.npc};
  const int lg = (int)(wg - E.start[IB_SWEEP]), ntx = (int)E.gx[IB_SWEEP], nty = (int)E.gy[IB_SWEEP];
  const double *sn = sn_all + 9 * b.npc;
  if (E.tile == 1)
    sweep_body<false, double, 32, 2 * SW_TY>(b, lg, ntx, nty, E.ks_sweep, ph, ityprk, u, ff, v, ldv, v, sn, c,
                                             c + 3 * b.nnn, out_d ? B : A, out_d ? A : B, sq);
  else
    sweep_body<false, double, SW_TX, SW_TY>(b, lg, ntx, nty, E.ks_sweep, ph, ityprk, u, ff, v, ldv, v, sn, c,
                                            c + 3 * b.nnn, out_d ? B : A, out_d ? A : B, sq);
}

__global__ void __launch_bounds__(TX *TY) k_impl_update_batch(const ImplEnt *__restrict__ tab, int nb, double *v,
                                                              long ldv, const double *__restrict__ d, int nvar) {
  const unsigned wg = xcd_linear();
  const ImplEnt &E = tab[tab_find(tab, nb, IB_CELL, wg)];
  update_body(E.b, tab_tile(E, IB_CELL, wg), v, ldv, d, nvar);
}

extern "C" int codens_implimf_3d_batch(int32_t nblk, const codens_block *blocks, const codens_phys *cp,
                                       const double *u, const double *dt, double *v, int64_t ldv, double *d,
                                       const double *ff, const double *mu, const double *mut, int32_t lmx,
                                       const double *sn_all, const double *vol, double dtpas, int32_t kfmg,
                                       int32_t ityprk, const double *cson, void *stream) {
  const char *what = "codens_implimf_3d_batch";
  if (nblk <= 0 || !blocks || !cp || !u || !dt || !v || !d || !mu || !mut || !sn_all || !vol || !cson ||
      (ityprk && !ff) || lmx < 0) {
    codens_set_error("%s: null or invalid argument", what);
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  std::vector<ImplEnt> ent(nblk);
  std::vector<int32_t> key;
  std::vector<std::pair<long, long>> span(nblk);
  unsigned tot[IB_N] = {0};
  for (int l = 0; l < nblk; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    const int32_t *w = (const int32_t *)&blocks[l];
    key.insert(key.end(), w, w + sizeof(codens_block) / sizeof(int32_t));
    Blk b = make_blk(&blocks[l]);
    span[l] = {b.npc, b.npc + b.nnn};
    ImplEnt &E = ent[l];
    E.b = b;
    dim3 cg = cell_grid(b);
    {
      std::lock_guard<std::mutex> lk(g_tile_mu);
      // the all-blocks launch holds shapes 0 and 1 (a forced 16 x 16 takes
      // the model's choice between them)
      E.tile = (g_tile_mode >= 0 && g_tile_mode < N_BATCH_TILES) ? g_tile_mode : tile_model(b);
    }
    const int tx = tile_tx(E.tile), ty = tile_ty(E.tile);
    unsigned ntx = (b.i2 - b.i1 + tx - 1) / tx, nty = (b.j2 - b.j1 + ty - 1) / ty;
    // full slabs: the launch has workgroups to spare, and these kernels give
    // the same values at any slab depth
    E.ks_coef = CK_KS;
    E.ks_sweep = SW_KS;
    const unsigned nkz = (unsigned)((b.k2 - b.k1 + E.ks_sweep - 1) / E.ks_sweep);
    unsigned g[IB_N][3] = {{(unsigned)((b.nnn + 255) / 256), 1, 1},
                           {cg.x, cg.y, cg.z},
                           {ntx, nty, nkz},
                           {cg.x, cg.y, (unsigned)((b.k2 - b.k1 + E.ks_coef - 1) / E.ks_coef)}};
    for (int p = 0; p < IB_N; p++) {
      E.start[p] = tot[p];
      E.gx[p] = g[p][0];
      E.gy[p] = g[p][1];
      tot[p] += g[p][0] * g[p][1] * g[p][2];
    }
  }
  std::sort(span.begin(), span.end());
  for (int l = 0; l < nblk; l++) {
    if ((l == 0 && span[0].first != 0) || (l > 0 && span[l].first < span[l - 1].second) || span[l].second > ldv) {
      codens_set_error("%s: the blocks must start at entry 0, not overlap, and lie within ldv", what);
      return CODENS_EINVAL;
    }
  }
  const long tot_n = span[nblk - 1].second;
  for (int l = 0; l < nblk; l++) key.push_back(ent[l].tile);
  ImplEnt *tab = (ImplEnt *)codens_device_table(1, key, ent.data(), sizeof(ImplEnt) * nblk, st, what);
  if (!tab) return CODENS_EHIP;
  double *ws = (double *)codens_workspace(sizeof(double) * 9 * tot_n, 2);
  if (!ws) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  Phys ph = make_phys(cp);
  const ImplEnt *ct = tab;
  const int nb = nblk;
  const double *fx = ityprk ? ff : u;
  hipLaunchKernelGGL(k_zero5_batch, dim3(tot[IB_ZERO]), dim3(256), 0, st, ct, nb, d, (long)ldv);
  {
    CodensTimed _tm("implimf_coef", st);
    hipLaunchKernelGGL(k_impl_coef_batch, dim3(tot[IB_COEF]), dim3(TX, TY), 0, st, ct, nb, ph, (const double *)v,
                       (long)ldv, mu, mut, dt, vol, cson, sn_all, dtpas, (int)kfmg, ws);
  }
  for (int s = 1; s <= lmx; s++) {
    int out_d = (lmx - s) % 2 == 0;
    if (s == 1) {
      CodensTimed _tm("implimf_sweep1", st);
      hipLaunchKernelGGL(k_impl_sweep1_batch, dim3(tot[IB_CELL]), dim3(TX, TY), 0, st, ct, nb, (int)ityprk, u, fx,
                         (long)ldv, ws, d, out_d);
    } else {
      CodensTimed _tm("implimf_sweep", st);
      hipLaunchKernelGGL(k_impl_sweep_batch, dim3(tot[IB_SWEEP]), dim3(SW_TX * SW_TY), 0, st, ct, nb, ph,
                         (int)ityprk, u, fx, (const double *)v, (long)ldv, sn_all, ws, d, out_d, s % 2 == 0);
    }
  }
  {
    CodensTimed _tm("implimf_update", st);
    hipLaunchKernelGGL(k_impl_update_batch, dim3(tot[IB_CELL]), dim3(TX, TY), 0, st, ct, nb, v, (long)ldv,
                       (const double *)d, 5);
  }
  return codens_check_launch(what);
}

// ---------------------------------------------------------------------------
// FP32 defect-correction form.  np planes of a source (stride lds, storage
// entries npc .. npc+nnn-1) rounded to float planes of stride nnn.
struct F32Src {
  const double *src;  // entry npc + m of plane e at src[e * lds + npc + m]
  long lds, npc;
  int np;
  float *dst;  // plane e at dst[e * nnn + m]
};
struct F32Set {
  F32Src s[5];
  int n;
};

// up to five such conversions in one launch (one pass over the block's
// entries)
__global__ void __launch_bounds__(256) k_to_f32_set(long nnn, F32Set set) {
  long m = blockIdx.x * 256L + threadIdx.x;
  if (m >= nnn) return;
  for (int q = 0; q < set.n; q++) {
    const F32Src &c = set.s[q];
    for (int e = 0; e < c.np; e++) c.dst[e * nnn + m] = (float)c.src[e * c.lds + c.npc + m];
  }
}

// d = the FP32 increment widened to FP64, v += d (real cells)
__global__ void __launch_bounds__(TX *TY) k_impl_update_f32(Blk b, double *v, long ldv, double *d,
                                                            const float *__restrict__ df) {
  CELL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
#pragma unroll
  for (int e = 0; e < 5; e++) {
    double x = (double)df[e * b.nnn + m];
    d[e * ldv + n] = x;
    v[e * ldv + n] = v[e * ldv + n] + x;
  }
}

extern "C" int codens_implimf_3d_f32(const codens_block *cb, const codens_phys *cp, const double *u,
                                     const double *dt, double *v, int64_t ldv, double *d, const double *ff,
                                     const double *mu, const double *mut, int32_t lmx, const double *sn,
                                     const double *vol, double dtpas, int32_t kfmg, int32_t ityprk, const double *ps,
                                     const double *cson, void *stream) {
  const char *what = "codens_implimf_3d_f32";
  int rc = codens_check_block(cb, what);
  if (rc) return rc;
  if (!cp || !u || !dt || !v || !d || !mu || !mut || !sn || !vol || !cson || (ityprk && !ff) || lmx < 0) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Phys ph = make_phys(cp);
  hipStream_t st = (hipStream_t)stream;
  const long nnn = b.nnn;
  double *ws = (double *)codens_workspace(sizeof(double) * 4 * nnn, 2);
  // float planes: v 5, u 5, ff 5, sn 9, coefe 3 + diag 1, two increments 5 + 5
  float *wf = (float *)codens_workspace(sizeof(float) * 38 * nnn, 6);
  if (!ws || !wf) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  float *vF = wf, *uF = wf + 5 * nnn, *fF = wf + 10 * nnn, *snF = wf + 15 * nnn, *cF = wf + 24 * nnn;
  float *dgF = cF + 3 * nnn, *dA = wf + 28 * nnn, *dB = wf + 33 * nnn;
  dim3 g1((unsigned)((nnn + 255) / 256)), t1(256);
  // d = 0 on the ghosts; the update writes every real cell
  if (lmx >= 1) codens_zero_shell(b, (long)b.npc, (long)(b.npc + nnn - 1), d, (long)ldv, 5, st);
  else hipLaunchKernelGGL(k_zero5, g1, t1, 0, st, d, (long)ldv, (long)b.npc, (long)(b.npc + nnn - 1));
  dim3 g = cell_grid(b), t(TX, TY);
  {
    CodensTimed _tm("implimf32_coef", st);
    const int ks = slab_depth((long)g.x * g.y, b.k2 - b.k1, CK_KS);
    dim3 gc(g.x, g.y, (unsigned)((b.k2 - b.k1 + ks - 1) / ks));
    hipLaunchKernelGGL(k_impl_coef_tiled, gc, t, 0, st, b, ph, v, (long)ldv, mu, mut, dt, vol, cson, sn, dtpas,
                       (int)kfmg, ws, ws + 3 * nnn, ks);
    F32Set set{};
    set.s[set.n++] = F32Src{(const double *)ws, nnn, 0L, 4, cF};
    set.s[set.n++] = F32Src{v, (long)ldv, (long)b.npc, 5, vF};
    set.s[set.n++] = F32Src{u, (long)ldv, (long)b.npc, 5, uF};
    if (ityprk) set.s[set.n++] = F32Src{ff, (long)ldv, (long)b.npc, 5, fF};
    set.s[set.n++] = F32Src{sn, nnn, 0L, 9, snF};
    hipLaunchKernelGGL(k_to_f32_set, g1, t1, 0, st, nnn, set);
  }
  // storage-indexed views of the block-local float planes (entry n at n - npc)
  const float *vS = vF - b.npc, *uS = uF - b.npc, *fS = fF - b.npc;
  TView<float> A = {dA, nnn, b.npc}, B = {dB, nnn, b.npc};
  TView<float> last = A;
  int tile = 0;
  // as in the FP64 phase: sub-iteration 1 from d = 0, then the tiled sweeps
  for (int s = 1; s <= lmx; s++) {
    TView<float> out = (s % 2) ? A : B, in = (s % 2) ? B : A;
    if (s == 1) {
      CodensTimed _tm("implimf32_sweep1", st);
      hipLaunchKernelGGL(k_impl_sweep1<float>, g, t, 0, st, b, (int)ityprk, uS, ityprk ? fS : uS, nnn, dgF, out);
    } else {
      if (s == 2)
        tile = pick_tile<false, float>(b, st, ph, (int)ityprk, uS, ityprk ? fS : uS, vS, nnn, (const double *)nullptr,
                                       snF, cF, dgF, in, out);
      CodensTimed _tm("implimf32_sweep", st);
      launch_sweep<false, float>(tile, b, st, ph, (int)ityprk, uS, ityprk ? fS : uS, vS, nnn, (const double *)nullptr,
                                 snF, cF, dgF, in, out, s % 2 == 0);
    }
    last = out;
  }
  if (lmx > 0) {
    CodensTimed _tm("implimf32_update", st);
    hipLaunchKernelGGL(k_impl_update_f32, g, t, 0, st, b, v, (long)ldv, d, (const float *)last.p);
  }
  return codens_check_launch(what);
}

extern "C" int codens_implimf_3d(const codens_block *cb, const codens_phys *cp, const double *u, const double *dt,
                                 double *v, int64_t ldv, double *d, const double *ff, const double *mu,
                                 const double *mut, int32_t lmx, const double *sn, const double *vol, double dtpas,
                                 int32_t kfmg, int32_t ityprk, const double *ps, const double *cson,
                                 void *stream) {
  // p of the real cells is re-derived from v with the zpres formula; ps unused
  return implimf_common(false, cb, cp, u, dt, v, ldv, 5, d, ff, mu, mut, lmx, sn, vol, dtpas, 0., kfmg, ityprk,
                        ps, cson, stream, "codens_implimf_3d");
}

extern "C" int codens_implimf_eu(const codens_block *cb, const codens_phys *cp, const double *u, const double *dt,
                                 double *v, int64_t ldv, int32_t nvar, double *d, const double *ff, int32_t lmx,
                                 const double *sn, const double *vol, double dt1min, int32_t kfmg, int32_t ityprk,
                                 const double *ps, const double *cson, void *stream) {
  return implimf_common(true, cb, cp, u, dt, v, ldv, nvar, d, ff, nullptr, nullptr, lmx, sn, vol, 1., dt1min, kfmg,
                        ityprk, ps, cson, stream, "codens_implimf_eu");
}
