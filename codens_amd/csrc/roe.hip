// Roe-MUSCL residual (sch_roe / sch_roe_euler) on gfx950.
//
// Passes per call:
//  0. primitive velocities u = (rho u)/rho, v, w of every cell of the block,
//     once (the flux passes read each cell up to 12 times);
//  1. (ilim=1 only) slope-ratio chains: one lane per grid line marches the
//     faces of its line in storage order, because the reference replaces each
//     slope difference by its ratio in place, so a face's ratio divides by the
//     ratio (not the difference) of the face before it (sch_roe.f90:161-194).
//  2. face fluxes: one thread per face, all faces of one direction per launch,
//     written to a per-direction face array H (amount the high-side cell
//     subtracts and the low-side cell adds).
//  3. gather: u(c) = sum over directions of H(hi face) - H(lo face), plus the
//     forcing ff -= u when ityprk != 0; ghosts of u are zeroed as the
//     reference does (sch_roe.f90:110-125).
// The dissipation |A| dW is evaluated as P (|Lambda| (Q dW)), the same
// matrices as the reference's P |Lambda| Q product with a quarter of the work.
// The Roe average forms sqrt(rho_L rho_R) as rho_L sqrt(rho_R/rho_L) and
// multiplies by reciprocals of rho_L, rho_R, |S|, c, 1 + sqrt(rho_R/rho_L),
// sqrt(rho_L rho_R) and gamma-1 where the reference divides; the limiter
// takes 1/r of its slope ratio and the fused kernel 1/rho of its cells the
// same way (rcp64; a product by it is at most two roundings from the
// reference's quotient).
#include "common.hpp"

#include <algorithm>
#include <utility>
#include <vector>

using namespace codens;

struct Prim {
  double r, u, v, w, p;
};

// primitive state of the cell at storage index n; uvw: the velocity planes of
// pass 0 (stride nnn, block-local index m = n - npc)
__device__ inline Prim prim_at(const double *__restrict__ v, const double *__restrict__ uvw, long nnn, long npc,
                               const double *__restrict__ ps, long n) {
  Prim q;
  long m = n - npc;
  q.r = v[n];
  q.u = uvw[m];
  q.v = uvw[nnn + m];
  q.w = uvw[2 * nnn + m];
  q.p = ps[n];
  return q;
}

// pass 0: uvw(e-1, m) = v(e, n) / v(1, n), e = 2..4, over the block's storage
__device__ __forceinline__ void prim_uvw_body(const Blk &b, unsigned bx, const double *__restrict__ v, long ldv,
                                              double *__restrict__ uvw) {
  long m = bx * 256L + threadIdx.x;
  if (m >= b.nnn) return;
  long n = b.npc + m;
  double r = v[n];
  uvw[m] = v[ldv + n] / r;
  uvw[b.nnn + m] = v[2 * ldv + n] / r;
  uvw[2 * b.nnn + m] = v[3 * ldv + n] / r;
}

__global__ void __launch_bounds__(256) k_prim_uvw(Blk b, const double *__restrict__ v, long ldv,
                                                  double *__restrict__ uvw) {
  prim_uvw_body(b, blockIdx.x, v, ldv, uvw);
}

__device__ inline double pget(const Prim &q, int e) {
  return e == 0 ? q.r : (e == 1 ? q.u : (e == 2 ? q.v : (e == 3 ? q.w : q.p)));
}

__device__ inline double minmod_phi(double a) { return fmax(0., fmin(1., a)); }

// minmod_phi(1 / r) without the division where its value is 0 or 1: 1/r
// lies in (0, 1) only for r > 1 (clamped below the overflow of 1/r, where
// the reciprocal of the clamp stands in for 0); r <= 1 gives 1 when r is
// +0 or positive, 0 when r is negative or -0, as the division would.
__device__ inline double minmod_phi_inv(double r) {
  return r > 1. ? rcp64(fmin(r, 1e300)) : (signbit(r) ? 0. : 1.);
}

// Roe flux (2x the numerical flux . S) of one face, sch_roe.f90:327-473.
// tl, tr: stresses and heat flux of the left/right cells (9 values each) or
// all zero for the Euler form.
__device__ void roe_face(const Phys &ph, double s0, double s1, double s2, const double L[5], const double R[5],
                         const double tl[9], const double tr[9], double f[5]) {
  const double gam = ph.gam, gam1 = ph.gam1, pinfl = ph.pinfl;
  double rhol = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  double rhor = R[0], ur = R[1], vr = R[2], wr = R[3], prr = R[4];
  const double igam1 = ph.igam1;
  double cnds = sqrt(s0 * s0 + s1 * s1 + s2 * s2), icn = rcp64(cnds);
  double nx = s0 * icn, ny = s1 * icn, nz = s2 * icn;
  double irl = rcp64(rhol), irr = rcp64(rhor);
  double al2 = gam * pl * irl, ar2 = gam * prr * irr;
  double q2l = ul * ul + vl * vl + wl * wl, q2r = ur * ur + vr * vr + wr * wr;
  double hl = al2 * igam1 + 0.5 * q2l, hr = ar2 * igam1 + 0.5 * q2r;
  double el = (pl * igam1 + pinfl) * irl + 0.5 * q2l;
  double er = (prr * igam1 + pinfl) * irr + 0.5 * q2r;
  double gd = sqrt(rhor * irl), gd1 = rcp64(1. + gd), gd2 = gd * gd1;
  double rhom = rhol * gd, rhomi = rcp64(rhom);  // sqrt(rho_L rho_R)
  double um = gd1 * ul + gd2 * ur, vm = gd1 * vl + gd2 * vr, wm = gd1 * wl + gd2 * wr;
  double hm = gd1 * hl + gd2 * hr;
  double vitm2 = 0.5 * (um * um + vm * vm + wm * wm);
  double am = sqrt(gam1 * (hm - vitm2)), ami = rcp64(am);
  double am2i = ami * ami;
  double vn = um * nx + vm * ny + wm * nz;
  double rhoiam = rhom * ami, rhoami = ami * rhomi;
  double l1 = fabs(vn * cnds), l4 = fabs(vn * cnds + am * cnds), l5 = fabs(vn * cnds - am * cnds);
  double dw0 = rhor - rhol;
  double dw1 = rhor * ur - rhol * ul;
  double dw2 = rhor * vr - rhol * vl;
  double dw3 = rhor * wr - rhol * wl;
  double dw4 = rhor * er - rhol * el;
  // a = Q dW (rows of Q as in sch_roe.f90:370-394)
  double g = gam1 * vitm2 * am2i, ga = gam1 * am2i;
  double qdw = gam1 * (um * dw1 + vm * dw2 + wm * dw3 - dw4);  // shared part of rows 1-3 (times am2i)
  double a0 = (1. - g) * nx * dw0 - (vm * nz - wm * ny) * rhomi * dw0 + nx * am2i * qdw +
              (nz * dw2 - ny * dw3) * rhomi;
  double a1 = (1. - g) * ny * dw0 - (wm * nx - um * nz) * rhomi * dw0 + ny * am2i * qdw +
              (nx * dw3 - nz * dw1) * rhomi;
  double a2 = (1. - g) * nz * dw0 - (um * ny - vm * nx) * rhomi * dw0 + nz * am2i * qdw +
              (ny * dw1 - nx * dw2) * rhomi;
  double mdw = (nx * dw1 + ny * dw2 + nz * dw3) * rhomi;
  double cdw = gam1 * rhoami * (vitm2 * dw0 - um * dw1 - vm * dw2 - wm * dw3 + dw4);
  double a3 = cdw - vn * rhomi * dw0 + mdw;
  double a4 = cdw + vn * rhomi * dw0 - mdw;
  (void)ga;
  a0 *= l1; a1 *= l1; a2 *= l1; a3 *= l4; a4 *= l5;
  // di = P (|Lambda| a), P as in sch_roe.f90:396-420
  double h = 0.5 * rhoiam;
  double di0 = nx * a0 + ny * a1 + nz * a2 + h * (a3 + a4);
  double di1 = um * nx * a0 + (um * ny - rhom * nz) * a1 + (um * nz + rhom * ny) * a2 +
               h * ((um + nx * am) * a3 + (um - nx * am) * a4);
  double di2 = (vm * nx + rhom * nz) * a0 + vm * ny * a1 + (vm * nz - rhom * nx) * a2 +
               h * ((vm + ny * am) * a3 + (vm - ny * am) * a4);
  double di3 = (wm * nx - rhom * ny) * a0 + (wm * ny + rhom * nx) * a1 + wm * nz * a2 +
               h * ((wm + nz * am) * a3 + (wm - nz * am) * a4);
  double di4 = (vitm2 * nx + rhom * (vm * nz - wm * ny)) * a0 + (vitm2 * ny + rhom * (wm * nx - um * nz)) * a1 +
               (vitm2 * nz + rhom * (um * ny - vm * nx)) * a2 + h * ((hm + am * vn) * a3 + (hm - am * vn) * a4);
  double dfxx = rhor * (ur * ur) + prr - pinfl - tr[0] + rhol * (ul * ul) + pl - pinfl - tl[0];
  double dfxy = rhor * ur * vr - tr[1] + rhol * ul * vl - tl[1];
  double dfxz = rhor * ur * wr - tr[2] + rhol * ul * wl - tl[2];
  double dfyy = rhor * (vr * vr) + prr - pinfl - tr[3] + rhol * (vl * vl) + pl - pinfl - tl[3];
  double dfyz = rhor * vr * wr - tr[4] + rhol * vl * wl - tl[4];
  double dfzz = rhor * (wr * wr) + prr - pinfl - tr[5] + rhol * (wl * wl) + pl - pinfl - tl[5];
  double Er = rhor * er + prr - pinfl, El = rhol * el + pl - pinfl;
  double dfex = Er * ur + El * ul - (tr[0] * ur + tr[1] * vr + tr[2] * wr) - (tl[0] * ul + tl[1] * vl + tl[2] * wl) -
                tr[6] - tl[6];
  double dfey = Er * vr + El * vl - (tr[1] * ur + tr[3] * vr + tr[4] * wr) - (tl[1] * ul + tl[3] * vl + tl[4] * wl) -
                tr[7] - tl[7];
  double dfez = Er * wr + El * wl - (tr[2] * ur + tr[4] * vr + tr[5] * wr) - (tl[2] * ul + tl[4] * vl + tl[5] * wl) -
                tr[8] - tl[8];
  f[0] = (rhor * ur + rhol * ul) * s0 + (rhor * vr + rhol * vl) * s1 + (rhor * wr + rhol * wl) * s2 - di0;
  f[1] = dfxx * s0 + dfxy * s1 + dfxz * s2 - di1;
  f[2] = dfxy * s0 + dfyy * s1 + dfyz * s2 - di2;
  f[3] = dfxz * s0 + dfyz * s1 + dfzz * s2 - di3;
  f[4] = dfex * s0 + dfey * s1 + dfez * s2 - di4;
}

// Half the Roe flux of one interior face between cells a-1 and a
// (sch_roe.f90:196-474): q[s][e], primitive variable e of cell a-2+s (s =
// 0..3); with lim, the MUSCL slopes are limited by the chained ratios rl of
// face a and rh of face a+1 (minmod, sch_roe.f90:1500-1504); cm = muscl/4;
// tl, tr: stresses and heat flux of cells a-1 and a (zero for Euler).
__device__ __forceinline__ void muscl_roe_half(const Phys &ph, double s0, double s1, double s2, const double q[4][5],
                                               bool lim, const double rl[5], const double rh[5], double cm, double xk,
                                               const double tl[9], const double tr[9], double f[5]) {
  double L[5], R[5];
  if (lim) {
#pragma unroll
    for (int e = 0; e < 5; e++) {
      double qa = q[0][e], qb = q[1][e], qc = q[2][e], qd = q[3][e];
      L[e] = qb + cm * minmod_phi(rl[e]) * (1. - xk) * (qb - qa) + cm * minmod_phi_inv(rl[e]) * (1. + xk) * (qc - qb);
      R[e] = qc - cm * minmod_phi(rh[e]) * (1. + xk) * (qc - qb) - cm * minmod_phi_inv(rh[e]) * (1. - xk) * (qd - qc);
    }
  } else {
#pragma unroll
    for (int e = 0; e < 5; e++) {
      double qa = q[0][e], qb = q[1][e], qc = q[2][e], qd = q[3][e];
      L[e] = qb + cm * ((1. - xk) * (qb - qa) + (1. + xk) * (qc - qb));
      R[e] = qc - cm * ((1. + xk) * (qc - qb) + (1. - xk) * (qd - qc));
    }
  }
  roe_face(ph, s0, s1, s2, L, R, tl, tr, f);
#pragma unroll
  for (int e = 0; e < 5; e++) f[e] *= 0.5;
}

// Chained slope ratios of the limiter (sch_roe.f90:146-194).  Along each grid
// line the reference first stores the differences r(a) = q(a) - q(a-1) of
// every face, then replaces r(a) by r(a)/r(a-1) in storage order, so the
// divisor is the ratio just written for the previous face (reset to 1 when
// |r(a-1) r(a) r(a+1)| <= tiny).  That recurrence is sequential along a line;
// the lines, the five variables and the three directions are independent, so
// one lane runs one (direction, line, variable) chain.  A workgroup is one
// wave: one variable of 64 lines, so the variable is uniform in the wave.
//
// Faces a = a1+1 .. a2-1 of a line get a ratio, stored at the face's storage
// index (cell a); the chain reads q at cells a1-1 .. a2.
constexpr int CH_LINES = 64;

// plane of primitive variable e (rho, u, v, w, p), indexed by storage index
__device__ inline const double *chain_q(const Blk &b, const double *v, const double *uvw, const double *ps, int e) {
  return e == 0 ? v : (e == 4 ? ps : uvw + (long)(e - 1) * b.nnn - b.npc);
}

struct ChainState {
  double prev, cur, qc;
  // one face: qn = q(a+1); returns the ratio stored for face a
  __device__ inline double step(double qn) {
    double nxt = qn - qc;
    double r = fabs(prev * cur * nxt) <= DBL_MIN ? 1. : cur / prev;
    prev = r;
    cur = nxt;
    qc = qn;
    return r;
  }
};

// Lines of direction d: line -> first face a1 storage index.  The line
// index runs fastest over the first transverse direction (i for d = 1, 2).
__device__ inline long chain_line_base(const Blk &b, int d, long line) {
  int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  int na = b.hi(ta) - b.lo(ta);
  int c[3];
  c[ta] = b.lo(ta) + (int)(line % na);
  c[tb] = b.lo(tb) + (int)(line / na);
  c[d] = b.lo(d);
  return b.idx(c[0], c[1], c[2]);
}

// Directions j and k: adjacent lanes hold lines adjacent in i, so every load
// and store of the wave is one run of 64 consecutive doubles.  Each lane
// keeps CH_B cells of its line in registers, with the loads of the next batch
// in flight while the current batch runs.  blockIdx.x: 64-line group,
// blockIdx.y: variable, blockIdx.z: direction (1 + z).
constexpr int CH_B = 8;

__device__ __forceinline__ void chain_jk_body(const Blk &b, dim3 bi, const double *__restrict__ v,
                                              const double *__restrict__ uvw, const double *__restrict__ ps,
                                              double *__restrict__ R) {
  const int d = 1 + bi.z, e = bi.y;
  int ta = 0, tb = d == 2 ? 1 : 2;
  long nl = (long)(b.hi(ta) - b.lo(ta)) * (b.hi(tb) - b.lo(tb));
  long line = (long)bi.x * CH_LINES + threadIdx.x;
  if (line >= nl) return;
  int a1 = b.lo(d), a2 = b.hi(d);
  long inc = b.inc(d);
  long n = chain_line_base(b, d, line);
  double *Re = R + (5L * d + e) * b.nnn - b.npc;
  const double *q = chain_q(b, v, uvw, ps, e);
  ChainState cs;
  double qm = q[n - inc], q0 = q[n];
  cs.qc = q[n + inc];
  cs.prev = q0 - qm;
  cs.cur = cs.qc - q0;
  double xa[CH_B], xb[CH_B];
  auto load_batch = [&](int a0, double *x) {
#pragma unroll
    for (int s = 0; s < CH_B; s++)  // unconditional, clamped to the line
      x[s] = q[n + (long)(min(a0 + s, a2 - 1) - a1 + 1) * inc];
  };
  auto run_batch = [&](int a0, const double *x) {
#pragma unroll
    for (int s = 0; s < CH_B; s++) {
      if (a0 + s > a2 - 1) break;
      Re[n + (long)(a0 + s - a1) * inc] = cs.step(x[s]);
    }
  };
  int a0 = a1 + 1;
  if (a0 <= a2 - 1) load_batch(a0, xa);
  while (a0 <= a2 - 1) {
    load_batch(a0 + CH_B, xb);
    __builtin_amdgcn_sched_barrier(0);
    run_batch(a0, xa);
    a0 += CH_B;
    if (a0 > a2 - 1) break;
    load_batch(a0 + CH_B, xa);
    __builtin_amdgcn_sched_barrier(0);
    run_batch(a0, xb);
    a0 += CH_B;
  }
}

__global__ void __launch_bounds__(64) k_limiter_chain_jk(Blk b, const double *__restrict__ v,
                                                         const double *__restrict__ uvw,
                                                         const double *__restrict__ ps, double *__restrict__ R) {
  chain_jk_body(b, dim3(blockIdx.x, blockIdx.y, blockIdx.z), v, uvw, ps, R);
}

// Direction i: a lane's line is contiguous in memory, so lanes walking their
// own lines would touch 64 cache lines per access.  The wave instead moves
// chunks of CH_C cells of its 64 lines through LDS: coalesced loads (each
// run of CH_C consecutive cells of one line) give q, the lanes run their
// chains over the chunk in LDS (ratios written in place), and coalesced
// stores write the ratios back.  blockIdx.x: 64-line group,
// blockIdx.y: variable.
constexpr int CH_C = 32;

__device__ __forceinline__ void chain_i_body(const Blk &b, dim3 bi, const double *__restrict__ v,
                                             const double *__restrict__ uvw, const double *__restrict__ ps,
                                             double *__restrict__ R) {
  __shared__ double qs[CH_LINES * (CH_C + 1)];
  __shared__ long base[CH_LINES];
  const int e = bi.y, lane = threadIdx.x;
  long nl = (long)(b.j2 - b.j1) * (b.k2 - b.k1);
  long line0 = (long)bi.x * CH_LINES;
  int nlines = (int)min((long)CH_LINES, nl - line0);
  int a1 = b.i1, a2 = b.i2;
  const double *q = chain_q(b, v, uvw, ps, e);
  // each lane's own line (lanes past the last line mirror the last one and
  // neither run a chain nor store)
  const bool own = lane < nlines;
  long n = chain_line_base(b, 0, line0 + min(lane, nlines - 1));
  base[lane] = n;
  ChainState cs;
  {
    double qm = q[n - 1], q0 = q[n];
    cs.qc = q[n + 1];
    cs.prev = q0 - qm;
    cs.cur = cs.qc - q0;
  }
  double *Re = R + (long)e * b.nnn - b.npc;
  __syncthreads();
  // chunk a0 (faces a0 .. a0+nc-1) needs q at cells a0+1 .. a0+nc of every
  // line; load pass it covers lines it*(64/CH_C) + lane/CH_C, cell
  // lane%CH_C, clamped to the chunk.  The next chunk's loads are issued as
  // soon as the current chunk is in LDS, so they are in flight while its
  // chains run and its ratios are stored.
  constexpr int LPP = 64 / CH_C, NPASS = CH_LINES / LPP;
  const int c = lane % CH_C, lsub = lane / CH_C;
  auto load_chunk = [&](int a0, double *qv) {
    const int nc = min(CH_C, a2 - a0), cc = min(c, nc - 1);
#pragma unroll
    for (int it = 0; it < NPASS; it++) {
      int l = min(it * LPP + lsub, nlines - 1);
      qv[it] = q[base[l] + (a0 + 1 + cc - a1)];
    }
  };
  auto run_chunk = [&](int a0, const double *qv, double *qn) {
    const int nc = min(CH_C, a2 - a0);
#pragma unroll
    for (int it = 0; it < NPASS; it++) qs[(it * LPP + lsub) * (CH_C + 1) + c] = qv[it];
    __syncthreads();
    if (a0 + CH_C <= a2 - 1) load_chunk(a0 + CH_C, qn);
    if (own) {
#pragma unroll
      for (int k = 0; k < CH_C; k++) {
        if (k >= nc) break;
        double *slot = &qs[lane * (CH_C + 1) + k];
        *slot = cs.step(*slot);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NPASS; it++) {
      int l = it * LPP + lsub;
      if (l < nlines && c < nc) Re[base[l] + (a0 + c - a1)] = qs[l * (CH_C + 1) + c];
    }
    __syncthreads();
  };
  double qa[NPASS], qb[NPASS];
  int a0 = a1 + 1;
  if (a0 <= a2 - 1) load_chunk(a0, qa);
  while (a0 <= a2 - 1) {
    run_chunk(a0, qa, qb);
    a0 += CH_C;
    if (a0 > a2 - 1) break;
    run_chunk(a0, qb, qa);
    a0 += CH_C;
  }
}

__global__ void __launch_bounds__(64) k_limiter_chain_i(Blk b, const double *__restrict__ v,
                                                        const double *__restrict__ uvw,
                                                        const double *__restrict__ ps, double *__restrict__ R) {
  chain_i_body(b, dim3(blockIdx.x, blockIdx.y, 1), v, uvw, ps, R);
}

// physical flux . S of a ghost state (sch_roe.f90:497-578)
__device__ inline void ghost_flux(const Phys &ph, const double w[5], double p, const double T[9], double rdiv,
                                  int euler, double s0, double s1, double s2, double f[5]) {
  double pinfl = ph.pinfl;
  double fxx = w[1] * (w[1] / w[0]) + p - pinfl - T[0];
  double fxy = w[2] * (w[1] / w[0]) - T[1];
  double fxz = w[3] * (w[1] / w[0]) - T[2];
  double fyy = w[2] * (w[2] / w[0]) + p - pinfl - T[3];
  double fyz = w[3] * (w[2] / w[0]) - T[4];
  double fzz = w[3] * (w[3] / w[0]) + p - pinfl - T[5];
  double fex, fey, fez;
  if (euler) {
    fex = (w[4] + p - pinfl) * w[1] / rdiv;
    fey = (w[4] + p - pinfl) * w[2] / rdiv;
    fez = (w[4] + p - pinfl) * w[3] / rdiv;
  } else {
    fex = ((w[4] + p - pinfl - T[0]) * w[1] - T[1] * w[2] - T[2] * w[3]) / rdiv - T[6];
    fey = ((w[4] + p - pinfl - T[3]) * w[2] - T[1] * w[1] - T[4] * w[3]) / rdiv - T[7];
    fez = ((w[4] + p - pinfl - T[5]) * w[3] - T[2] * w[1] - T[4] * w[2]) / rdiv - T[8];
  }
  f[0] = w[1] * s0 + w[2] * s1 + w[3] * s2;
  f[1] = fxx * s0 + fxy * s1 + fxz * s2;
  f[2] = fxy * s0 + fyy * s1 + fyz * s2;
  f[3] = fxz * s0 + fyz * s1 + fzz * s2;
  f[4] = fex * s0 + fey * s1 + fez * s2;
}

// Minimum waves per SIMD the face-flux kernel is compiled for (its register
// budget: 3 -> up to 168 VGPRs, 4 -> 128).
#ifndef CODENS_FF_WAVES
#define CODENS_FF_WAVES 3
#endif

// ---------------------------------------------------------------------------
// pass 2: the half face fluxes H of direction d, one thread per face: the
// physical flux of the ghost state on the two boundary faces
// (sch_roe.f90:497-578), else 0.5 x the Roe flux of the MUSCL states
// (sch_roe.f90:196-474).  H[e*nnn + m] holds the face whose storage index is
// m (cells m-inc | m).
__device__ __forceinline__ void face_flux_body(const Blk &b, dim3 xb, const Phys &ph, int d, int ns, int ilim,
                                               double muscl, double xk, const double *__restrict__ v, long ldv,
                                               const double *__restrict__ ps, const double *__restrict__ tq,
                                               const double *__restrict__ sn, const double *__restrict__ uvw,
                                               const double *__restrict__ rr, double *__restrict__ H) {
  int c[3];
  c[0] = b.i1 + xb.x * TX + threadIdx.x;
  c[1] = b.j1 + xb.y * TY + threadIdx.y;
  c[2] = b.k1 + xb.z;
  int hi0 = b.i2 - 1 + (d == 0), hi1 = b.j2 - 1 + (d == 1);
  if (c[0] > hi0 || c[1] > hi1) return;
  int a = c[d], a1 = b.lo(d), a2 = b.hi(d);
  long inc = b.inc(d);
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double s0 = snv(sn, b.nnn, m, d, 0), s1 = snv(sn, b.nnn, m, d, 1), s2 = snv(sn, b.nnn, m, d, 2);
  double f[5];
  if (a == a1 || a == a2) {
    long ng = a == a1 ? n - inc : n;
    double w[5], T[9];
#pragma unroll
    for (int e = 0; e < 5; e++) w[e] = v[e * ldv + ng];
#pragma unroll
    for (int q = 0; q < 9; q++) T[q] = ns ? tq[q * ldv + ng] : 0.;
    double rdiv = (!ns && a == a1 && d > 0) ? v[n] : w[0];  // sch_roe_euler.f90:630-632, 912-914
    ghost_flux(ph, w, ps[ng], T, rdiv, !ns, s0, s1, s2, f);
  } else {
    double q[4][5];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      Prim c = prim_at(v, uvw, b.nnn, b.npc, ps, n + (s - 2) * inc);
#pragma unroll
      for (int e = 0; e < 5; e++) q[s][e] = pget(c, e);
    }
    const bool lim = ilim == 1 && a <= a2 - 2;
    double rl[5], rh[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      rl[e] = lim ? rr[e * b.nnn + m] : 0.;
      rh[e] = lim ? rr[e * b.nnn + m + inc] : 0.;
    }
    double tl[9], tr[9];
#pragma unroll
    for (int q2 = 0; q2 < 9; q2++) {
      tl[q2] = ns ? tq[q2 * ldv + n - inc] : 0.;
      tr[q2] = ns ? tq[q2 * ldv + n] : 0.;
    }
    muscl_roe_half(ph, s0, s1, s2, q, lim, rl, rh, 0.25 * muscl, xk, tl, tr, f);
  }
#pragma unroll
  for (int e = 0; e < 5; e++) H[e * b.nnn + m] = f[e];
}

__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(CODENS_FF_WAVES)))
k_face_flux(Blk b, Phys ph, int d, int ns, int ilim, double muscl,
                                                      double xk, const double *__restrict__ v, long ldv,
                                                      const double *__restrict__ ps,
                                                      const double *__restrict__ tq,
                                                      const double *__restrict__ sn,
                                                      const double *__restrict__ uvw,
                                                      const double *__restrict__ rr, double *__restrict__ H) {
  // the j pass runs in reverse order, so that each pass begins on the
  // cells (v, p, uvw, tq) the previous one read last
  face_flux_body(b, xcd_block(d == 1), ph, d, ns, ilim, muscl, xk, v, ldv, ps, tq, sn, uvw, rr, H);
}

// ---------------------------------------------------------------------------
// pass 3: gather over the zeroed range indc(i1-1,j1-1,k1-1)..indc(i2,j2,k2)
// (sch_roe.f90:113-125, 448-474 regrouped per cell):
// u = (H_i(hi) - H_i(lo)) + (H_j(hi) - H_j(lo)) + (H_k(hi) - H_k(lo)) on
// real cells, 0 elsewhere; ff -= u when ityprk != 0.
__device__ __forceinline__ void gather_body(const Blk &b, unsigned bx, long n0, long n1, const double *__restrict__ H,
                                            double *u, long ldv, double *ff, int ityprk) {
  long n = n0 + bx * 256L + threadIdx.x;
  if (n > n1) return;
  long m = n - b.npc;
  int i = b.id1 + (int)(m % b.nid);
  int j = b.jd1 + (int)((m / b.nid) % (b.nijd / b.nid));
  int k = b.kd1 + (int)(m / b.nijd);
  const double *H0 = H, *H1 = H + 5 * b.nnn, *H2 = H + 10 * b.nnn;
  if (!b.real(i, j, k)) {
#pragma unroll
    for (int e = 0; e < 5; e++) u[e * ldv + n] = 0.;
    return;
  }
#pragma unroll
  for (int e = 0; e < 5; e++) {
    long o = e * b.nnn + m;
    double r = H0[o + 1] - H0[o];
    r = r + (H1[o + b.nid] - H1[o]);
    r = r + (H2[o + b.nijd] - H2[o]);
    u[e * ldv + n] = r;
    if (ityprk) ff[e * ldv + n] = ff[e * ldv + n] - r;
  }
}

__global__ void __launch_bounds__(256) k_gather(Blk b, long n0, long n1, const double *__restrict__ H, double *u,
                                                long ldv, double *ff, int ityprk) {
  gather_body(b, gridDim.x - 1 - blockIdx.x, n0, n1, H, u, ldv, ff, ityprk);  // after the k pass, in reverse
}

// ---------------------------------------------------------------------------
// Fused residual: the face fluxes and the gather of passes 2-3 in one kernel,
// with no half-flux planes H in HBM.  A workgroup owns a tile of FT x FT
// cells in i-j and marches a slab of up to FS k-planes.  Per plane it stages
// the primitive variables of the tile with a 2-cell halo in i and j (for the
// MUSCL stencil) and the stresses/heat flux with a 1-cell halo in LDS,
// evaluates every i face (FT+1 per row) and j face (FT+1 per column) of the
// tile once, and keeps each thread's own k column in registers: the face
// between planes k and k+1 is evaluated once per cell and carried to the
// next plane as its low face.  A cell's residual then sums the same face
// values in the same order as the gather,
//   u = (F_i(hi) - F_i(lo)) + (F_j(hi) - F_j(lo)) + (F_k(hi) - F_k(lo)),
// so each face value, and u, are those of passes 2-3 (each face is the same
// function of the same operands).  The k face between two slabs is evaluated
// by both (one extra face per slab and cell).  A slab is FS planes deep on
// large blocks, shallower on small ones (slab_depth), in the per-block and
// the all-blocks launch alike.
//
// Faces on the block's boundary (a = a1, a2) are the physical flux of the
// ghost state (sch_roe.f90:497-578), read from HBM (they are few).  The
// FT+1-th face of each row and column (32 per plane) is taken by half a wave,
// a different wave on each plane, so the extra work is spread over the four.
constexpr int FT = 16;       // tile cells per direction (i, j)
constexpr int FPP = FT + 4;  // pitch of the primitive plane (halo 2)
constexpr int FQP = FT + 2;  // pitch of the stress plane (halo 1)
constexpr int FS = 8;        // k planes per workgroup, at most

struct FusedArgs {
  Blk b;
  Phys ph;
  double cm, xk;
  const double *v;
  long ldv;
  const double *ps, *tq, *sn, *rr;
  double *u, *ff;
  int ityprk;
  int ks;  // k planes per workgroup
};

// primitive variables of storage entry n: rho, u, v, w (the divisions of
// pass 0) and p
__device__ __forceinline__ void fused_prim(const FusedArgs &A, long n, double q[5]) {
  double r = A.v[n], ir = rcp64(r);
  q[0] = r;
  q[1] = A.v[A.ldv + n] * ir;
  q[2] = A.v[2 * A.ldv + n] * ir;
  q[3] = A.v[3 * A.ldv + n] * ir;
  q[4] = A.ps[n];
}

// physical flux of the ghost state of boundary face (direction d, storage
// entry n of the face's high-side cell; ghost at n - inc on the low face, n
// on the high one)
template <bool NS>
__device__ __forceinline__ void fused_ghost(const FusedArgs &A, int d, bool low, long n, long inc, double s0,
                                            double s1, double s2, double f[5]) {
  long ng = low ? n - inc : n;
  double w[5], T[9];
#pragma unroll
  for (int e = 0; e < 5; e++) w[e] = A.v[e * A.ldv + ng];
#pragma unroll
  for (int q = 0; q < 9; q++) T[q] = NS ? A.tq[q * A.ldv + ng] : 0.;
  double rdiv = (!NS && low && d > 0) ? A.v[n] : w[0];  // sch_roe_euler.f90:630-632, 912-914
  ghost_flux(A.ph, w, A.ps[ng], T, rdiv, !NS, s0, s1, s2, f);
}

template <bool NS, bool LIM>
__device__ __forceinline__ void fused_body(const FusedArgs &A, dim3 t) {
  const Blk &b = A.b;
  __shared__ double P[5][FPP * FPP];
  __shared__ double Q[NS ? 9 : 1][NS ? FQP * FQP : 1];
  __shared__ double Fi[5][(FT + 1) * FT];
  __shared__ double Fj[5][FT * (FT + 1)];
  const int tid = threadIdx.x, tx = tid % FT, ty = tid / FT;
  const int i0 = b.i1 + (int)t.x * FT, j0 = b.j1 + (int)t.y * FT;
  const int kb = b.k1 + (int)t.z * A.ks, ke = min(kb + A.ks, b.k2);
  const int njd = (int)(b.nijd / b.nid), nkd = (int)(b.nnn / b.nijd);
  const int id2 = b.id1 + (int)b.nid - 1, jd2 = b.jd1 + njd - 1, kd2 = b.kd1 + nkd - 1;
  auto cl = [](int x, int lo, int hi) { return min(max(x, lo), hi); };
  // own column (clamped into the block's storage: cells past the real ones
  // are the ghost layer, or beyond it, whose values no needed face uses)
  const int i = i0 + tx, j = j0 + ty;
  const bool cell = i < b.i2 && j < b.j2;
  const long col = b.idx(cl(i, b.id1, id2), cl(j, b.jd1, jd2), b.kd1);  // entry at k = kd1
  auto at = [&](int k) { return col + (long)(cl(k, b.kd1, kd2) - b.kd1) * b.nijd; };
  // halo element of this thread: prims (tid < 128) or stresses (128..191)
  int hii = 0, hjj = 0, hkind = 0;  // kind 1: prim, 2: stress
  if (tid < 64) {
    int side = tid / 32, r = tid % 32;
    hjj = r / 2; hii = side ? FT + (r % 2) : -2 + (r % 2); hkind = 1;
  } else if (tid < 128) {
    int h = tid - 64, side = h / 32, r = h % 32;
    hii = r % FT; hjj = side ? FT + r / FT : -2 + r / FT; hkind = 1;
  } else if (NS && tid < 192) {
    int g = tid - 128;
    if (g < 32) { hjj = g % FT; hii = (g / FT) ? FT : -1; }
    else { g -= 32; hii = g % FT; hjj = (g / FT) ? FT : -1; }
    hkind = 2;
  }
  const long hcol = b.idx(cl(i0 + hii, b.id1, id2), cl(j0 + hjj, b.jd1, jd2), b.kd1);
  const int so = (ty + 2) * FPP + tx + 2, sq = (ty + 1) * FQP + tx + 1;
  const double z9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // own column registers: c0 = q(k-1), cq = q(k), c2 = q(k+1); Hlo = the
  // low k face of plane k
  double c0[5], cq[5], c2[5], Hlo[5];
  fused_prim(A, at(kb - 1), c0);
  fused_prim(A, at(kb), cq);
  fused_prim(A, at(kb + 1), c2);
  // the low face of the slab's first plane
  if (cell) {
    const long n = b.idx(i, j, kb), m = n - b.npc;
    double s0 = snv(A.sn, b.nnn, m, 2, 0), s1 = snv(A.sn, b.nnn, m, 2, 1), s2 = snv(A.sn, b.nnn, m, 2, 2);
    if (kb == b.k1) {
      fused_ghost<NS>(A, 2, true, n, b.nijd, s0, s1, s2, Hlo);
    } else {
      double q[4][5], rl[5], rh[5], tl[9], tr[9];
      fused_prim(A, at(kb - 2), q[0]);
#pragma unroll
      for (int e = 0; e < 5; e++) { q[1][e] = c0[e]; q[2][e] = cq[e]; q[3][e] = c2[e]; }
      const long nl = at(kb - 1), nr = at(kb);
#pragma unroll
      for (int x = 0; x < 9; x++) {
        tl[x] = NS ? A.tq[x * A.ldv + nl] : 0.;
        tr[x] = NS ? A.tq[x * A.ldv + nr] : 0.;
      }
      const bool lim = LIM && kb <= b.k2 - 2;
#pragma unroll
      for (int e = 0; e < 5; e++) {
        rl[e] = lim ? A.rr[(10 + e) * b.nnn + m] : 0.;
        rh[e] = lim ? A.rr[(10 + e) * b.nnn + m + b.nijd] : 0.;
      }
      muscl_roe_half(A.ph, s0, s1, s2, q, lim, rl, rh, A.cm, A.xk, tl, tr, Hlo);
    }
  }
  for (int k = kb; k < ke; k++) {
    // plane k into LDS: own cell, own stresses, halo element
    {
      const long nk = at(k);
#pragma unroll
      for (int e = 0; e < 5; e++) P[e][so] = cq[e];
      if (NS) {
#pragma unroll
        for (int x = 0; x < 9; x++) Q[x][sq] = A.tq[x * A.ldv + nk];
      }
      const long nh = hcol + (long)(cl(k, b.kd1, kd2) - b.kd1) * b.nijd;
      if (hkind == 1) {
        double hv[5];
        fused_prim(A, nh, hv);
        const int sh = (hjj + 2) * FPP + hii + 2;
#pragma unroll
        for (int e = 0; e < 5; e++) P[e][sh] = hv[e];
      } else if (NS && hkind == 2) {
        const int sh = (hjj + 1) * FQP + hii + 1;
#pragma unroll
        for (int x = 0; x < 9; x++) Q[x][sh] = A.tq[x * A.ldv + nh];
      }
    }
    __syncthreads();
    // i and j faces of plane k; the extra (FT+1-th) faces by half of wave
    // (k - kb) % 4
    const int wave = tid / 64, lane = tid % 64;
    const int njob = (wave == (k - kb) % 4 && lane < 2 * FT) ? 3 : 2;
    for (int job = 0; job < njob; job++) {
      int d, fx, fy;
      if (job == 0) { d = 0; fx = tx; fy = ty; }
      else if (job == 1) { d = 1; fx = tx; fy = ty; }
      else if (lane < FT) { d = 0; fx = FT; fy = lane; }
      else { d = 1; fx = lane - FT; fy = FT; }
      const int fi = i0 + fx, fj = j0 + fy;
      const bool need = d == 0 ? (fi <= b.i2 && fj < b.j2) : (fi < b.i2 && fj <= b.j2);
      if (!need) continue;
      const int a = d == 0 ? fi : fj, a1 = d == 0 ? b.i1 : b.j1, a2 = d == 0 ? b.i2 : b.j2;
      const long inc = d == 0 ? 1 : b.nid;
      const long n = b.idx(fi, fj, k), m = n - b.npc;
      double s0 = snv(A.sn, b.nnn, m, d, 0), s1 = snv(A.sn, b.nnn, m, d, 1), s2 = snv(A.sn, b.nnn, m, d, 2);
      double f[5];
      if (a == a1 || a == a2) {
        fused_ghost<NS>(A, d, a == a1, n, inc, s0, s1, s2, f);
      } else {
        double q[4][5], rl[5], rh[5], tl[9], tr[9];
        const int dx = d == 0, dy = d == 1;
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int sl = (fy + (s - 2) * dy + 2) * FPP + fx + (s - 2) * dx + 2;
#pragma unroll
          for (int e = 0; e < 5; e++) q[s][e] = P[e][sl];
        }
        const int ql = (fy - dy + 1) * FQP + fx - dx + 1, qr = (fy + 1) * FQP + fx + 1;
#pragma unroll
        for (int x = 0; x < 9; x++) {
          tl[x] = NS ? Q[x][ql] : 0.;
          tr[x] = NS ? Q[x][qr] : 0.;
        }
        const bool lim = LIM && a <= a2 - 2;
#pragma unroll
        for (int e = 0; e < 5; e++) {
          rl[e] = lim ? A.rr[(5 * d + e) * b.nnn + m] : 0.;
          rh[e] = lim ? A.rr[(5 * d + e) * b.nnn + m + inc] : 0.;
        }
        muscl_roe_half(A.ph, s0, s1, s2, q, lim, rl, rh, A.cm, A.xk, tl, tr, f);
      }
      if (d == 0) {
#pragma unroll
        for (int e = 0; e < 5; e++) Fi[e][fy * (FT + 1) + fx] = f[e];
      } else {
#pragma unroll
        for (int e = 0; e < 5; e++) Fj[e][fy * FT + fx] = f[e];
      }
    }
    // the k face between planes k and k+1 of the own column
    double Hhi[5], c3[5];
    fused_prim(A, at(k + 2), c3);
    if (cell) {
      const long n = b.idx(i, j, k + 1), m = n - b.npc;
      double s0 = snv(A.sn, b.nnn, m, 2, 0), s1 = snv(A.sn, b.nnn, m, 2, 1), s2 = snv(A.sn, b.nnn, m, 2, 2);
      if (k + 1 == b.k2) {
        fused_ghost<NS>(A, 2, false, n, b.nijd, s0, s1, s2, Hhi);
      } else {
        double q[4][5], rl[5], rh[5], tl[9], tr[9];
#pragma unroll
        for (int e = 0; e < 5; e++) { q[0][e] = c0[e]; q[1][e] = cq[e]; q[2][e] = c2[e]; q[3][e] = c3[e]; }
        const long nr = at(k + 1);
#pragma unroll
        for (int x = 0; x < 9; x++) {
          tl[x] = NS ? Q[x][sq] : 0.;
          tr[x] = NS ? A.tq[x * A.ldv + nr] : 0.;
        }
        const bool lim = LIM && k + 1 <= b.k2 - 2;
#pragma unroll
        for (int e = 0; e < 5; e++) {
          rl[e] = lim ? A.rr[(10 + e) * b.nnn + m] : 0.;
          rh[e] = lim ? A.rr[(10 + e) * b.nnn + m + b.nijd] : 0.;
        }
        muscl_roe_half(A.ph, s0, s1, s2, q, lim, rl, rh, A.cm, A.xk, tl, tr, Hhi);
      }
    }
    __syncthreads();
    if (cell) {
      const long n = b.idx(i, j, k);
#pragma unroll
      for (int e = 0; e < 5; e++) {
        double r = Fi[e][ty * (FT + 1) + tx + 1] - Fi[e][ty * (FT + 1) + tx];
        r = r + (Fj[e][(ty + 1) * FT + tx] - Fj[e][ty * FT + tx]);
        r = r + (Hhi[e] - Hlo[e]);
        A.u[e * A.ldv + n] = r;
        if (A.ityprk) A.ff[e * A.ldv + n] = A.ff[e * A.ldv + n] - r;
        Hlo[e] = Hhi[e];
      }
    }
#pragma unroll
    for (int e = 0; e < 5; e++) { c0[e] = cq[e]; cq[e] = c2[e]; c2[e] = c3[e]; }
  }
}

template <bool NS, bool LIM>
__global__ void __launch_bounds__(FT *FT) __attribute__((amdgpu_waves_per_eu(2))) k_residual_fused(FusedArgs A) {
  fused_body<NS, LIM>(A, xcd_block());
}

// u = 0 on the entries of indc(i1-1,j1-1,k1-1)..indc(i2,j2,k2) that are not
// real cells (sch_roe.f90:110-125; the fused kernel writes the real ones)
__device__ __forceinline__ void zero_ghost_body(const Blk &b, unsigned bx, long n0, long n1, double *u, long ldv) {
  long n = n0 + bx * 256L + threadIdx.x;
  if (n > n1) return;
  long m = n - b.npc;
  int i = b.id1 + (int)(m % b.nid);
  int j = b.jd1 + (int)((m / b.nid) % (b.nijd / b.nid));
  int k = b.kd1 + (int)(m / b.nijd);
  if (b.real(i, j, k)) return;
#pragma unroll
  for (int e = 0; e < 5; e++) u[e * ldv + n] = 0.;
}


inline dim3 fused_grid(const Blk &b, int ks) {
  int nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  return dim3((nx + FT - 1) / FT, (ny + FT - 1) / FT, (nz + ks - 1) / ks);
}

// Residual path: 1 (default) the fused kernel, 0 the face passes + gather,
// 2 the fused kernel with the all-blocks launch persistent over a tile queue.
static int g_residual_fused = 1;

extern "C" int codens_residual_path(int32_t fused, int32_t *previous) {
  if (fused < 0 || fused > 2) {
    codens_set_error("codens_residual_path: 0 (face passes + gather), 1 (fused) or 2 (fused, persistent all-blocks "
                     "launch), not %d", fused);
    return CODENS_EINVAL;
  }
  if (previous) *previous = g_residual_fused;
  g_residual_fused = fused;
  return CODENS_OK;
}

static int roe_common(const codens_block *cb, const codens_phys *cp, const codens_scheme *sc, int ns,
                      int ityprk, double *u, const double *v, long ldv, double *ff, const double *tq,
                      const double *sn, const double *ps, hipStream_t st, const char *what) {
  int rc = codens_check_block(cb, what);
  if (rc) return rc;
  if (!cp || !sc || !u || !v || !sn || !ps || (ns && !tq) || (ityprk && !ff)) {
    codens_set_error("%s: null argument", what);
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  Phys ph = make_phys(cp);
  int ilim = ns ? sc->ilim : 0;
  const bool fused = g_residual_fused >= 1;
  double *H = fused ? nullptr : (double *)codens_workspace(sizeof(double) * 15 * b.nnn, 0);
  double *R = ilim == 1 ? (double *)codens_workspace(sizeof(double) * 15 * b.nnn, 1) : nullptr;
  const bool need_uvw = !fused || ilim == 1;  // the chains read u, v, w planes
  double *uvw = need_uvw ? (double *)codens_workspace(sizeof(double) * 3 * b.nnn, 4) : nullptr;
  if ((!fused && !H) || (need_uvw && !uvw) || (ilim == 1 && !R)) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  if (need_uvw) {
    CodensTimed _tm("roe_prim", st);
    hipLaunchKernelGGL(k_prim_uvw, dim3((unsigned)((b.nnn + 255) / 256)), dim3(256), 0, st, b, v, ldv, uvw);
  }
  if (ilim == 1) {
    CodensTimed _tm("roe_limiter_chain", st);
    long nl0 = (long)(b.j2 - b.j1) * (b.k2 - b.k1);
    long nl12 = (long)(b.i2 - b.i1) * max(b.k2 - b.k1, b.j2 - b.j1);
    hipLaunchKernelGGL(k_limiter_chain_i, dim3((unsigned)((nl0 + CH_LINES - 1) / CH_LINES), 5), dim3(64), 0, st, b,
                       v, uvw, ps, R);
    hipLaunchKernelGGL(k_limiter_chain_jk, dim3((unsigned)((nl12 + CH_LINES - 1) / CH_LINES), 5, 2), dim3(64), 0, st,
                       b, v, uvw, ps, R);
  }
  long n0 = b.idx(b.i1 - 1, b.j1 - 1, b.k1 - 1), n1 = b.idx(b.i2, b.j2, b.k2);
  if (fused) {
    dim3 g = fused_grid(b, FS);
    const int ks = slab_depth((long)g.x * g.y, b.k2 - b.k1, FS);
    g = fused_grid(b, ks);
    FusedArgs A{b, ph, 0.25 * sc->muscl, sc->xk, v, ldv, ps, ns ? tq : v, sn, R, u, ff, ityprk, ks};
    {
      CodensTimed _tm("roe_fused", st);
      if (ns && ilim == 1) hipLaunchKernelGGL((k_residual_fused<true, true>), g, dim3(FT * FT), 0, st, A);
      else if (ns) hipLaunchKernelGGL((k_residual_fused<true, false>), g, dim3(FT * FT), 0, st, A);
      else hipLaunchKernelGGL((k_residual_fused<false, false>), g, dim3(FT * FT), 0, st, A);
    }
    CodensTimed _tm("roe_zero_ghost", st);
    codens_zero_shell(b, n0, n1, u, ldv, 5, st);
    return codens_check_launch(what);
  }
  for (int d = 0; d < 3; d++) {
    int nx = b.i2 - b.i1 + (d == 0), ny = b.j2 - b.j1 + (d == 1), nz = b.k2 - b.k1 + (d == 2);
    dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz);
    CodensTimed _tm("roe_face_flux", st);
    hipLaunchKernelGGL(k_face_flux, g, dim3(TX, TY), 0, st, b, ph, d, ns, ilim, (double)sc->muscl, sc->xk, v, ldv,
                       ps, ns ? tq : v, sn, uvw, ilim == 1 ? R + 5 * b.nnn * d : nullptr, H + 5 * b.nnn * d);
  }
  {
    CodensTimed _tm("roe_gather", st);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((n1 - n0 + 256) / 256)), dim3(256), 0, st, b, n0, n1, H, u, ldv,
                       ff, ityprk);
  }
  return codens_check_launch(what);
}

extern "C" int codens_sch_roe_euler(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                                    int32_t ityprk, double *u, const double *v, int64_t ldv, double *ff,
                                    const double *sn, const double *ps, void *stream) {
  return roe_common(b, ph, sc, 0, ityprk, u, v, (long)ldv, ff, nullptr, sn, ps, (hipStream_t)stream,
                    "codens_sch_roe_euler");
}

extern "C" int codens_sch_roe(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                              int32_t ityprk, double *u, const double *v, int64_t ldv, double *ff,
                              const double *tq, const double *sn, const double *ps, void *stream) {
  return roe_common(b, ph, sc, 1, ityprk, u, v, (long)ldv, ff, tq, sn, ps, (hipStream_t)stream,
                    "codens_sch_roe");
}

// ---------------------------------------------------------------------------
// All blocks of a layout in one launch per pass (the residual of a
// multi-block stage in seven launches instead of seven per block).  A device
// table gives every block its extents and, for each pass, the first
// workgroup of the block in the flattened launch and the block's own grid
// (x, y; z follows); a workgroup finds its block by bisection over the
// first-workgroup column and runs the pass's body on the block's tile.
// The passes' workspaces are the per-block ones laid end to end: block l's
// part starts at its npc (times the planes).
enum { BP_PRIM, BP_CHAIN_I, BP_CHAIN_JK, BP_FACE0, BP_FACE1, BP_FACE2, BP_GATHER, BP_FUSED, BP_N };

struct BatchEnt {
  Blk b;
  long n0, n1;                  // gather range
  unsigned start[BP_N];         // first workgroup of this block, per pass
  unsigned gx[BP_N], gy[BP_N];  // this block's grid, per pass
  int ks;                       // k planes per workgroup of the fused pass
};

__global__ void __launch_bounds__(256) k_prim_uvw_batch(const BatchEnt *__restrict__ tab, int nb,
                                                        const double *__restrict__ v, long ldv, double *uvw) {
  const BatchEnt &E = tab[tab_find(tab, nb, BP_PRIM, blockIdx.x)];
  prim_uvw_body(E.b, blockIdx.x - E.start[BP_PRIM], v, ldv, uvw + 3 * E.b.npc);
}

__global__ void __launch_bounds__(64) k_chain_i_batch(const BatchEnt *__restrict__ tab, int nb,
                                                      const double *__restrict__ v, const double *__restrict__ uvw,
                                                      const double *__restrict__ ps, double *R) {
  const BatchEnt &E = tab[tab_find(tab, nb, BP_CHAIN_I, blockIdx.x)];
  chain_i_body(E.b, tab_tile(E, BP_CHAIN_I, blockIdx.x), v, uvw + 3 * E.b.npc, ps, R + 15 * E.b.npc);
}

__global__ void __launch_bounds__(64) k_chain_jk_batch(const BatchEnt *__restrict__ tab, int nb,
                                                       const double *__restrict__ v, const double *__restrict__ uvw,
                                                       const double *__restrict__ ps, double *R) {
  const BatchEnt &E = tab[tab_find(tab, nb, BP_CHAIN_JK, blockIdx.x)];
  chain_jk_body(E.b, tab_tile(E, BP_CHAIN_JK, blockIdx.x), v, uvw + 3 * E.b.npc, ps, R + 15 * E.b.npc);
}

__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(CODENS_FF_WAVES)))
k_face_flux_batch(const BatchEnt *__restrict__ tab, int nb, Phys ph, int d, int ns, int ilim, double muscl, double xk,
                  const double *__restrict__ v, long ldv, const double *__restrict__ ps,
                  const double *__restrict__ tq, const double *__restrict__ sn_all, const double *__restrict__ uvw,
                  const double *__restrict__ rr, double *H) {
  const unsigned wg = xcd_linear();
  const int pass = BP_FACE0 + d;
  const BatchEnt &E = tab[tab_find(tab, nb, pass, wg)];
  const long np = E.b.npc;
  face_flux_body(E.b, tab_tile(E, pass, wg), ph, d, ns, ilim, muscl, xk, v, ldv, ps, tq, sn_all + 9 * np,
                 uvw + 3 * np, ilim == 1 ? rr + 15 * np + 5 * E.b.nnn * d : nullptr, H + 15 * np + 5 * E.b.nnn * d);
}

// the fused residual of every block: workgroup -> block (bisection) -> tile
template <bool NS, bool LIM>
__global__ void __launch_bounds__(FT *FT) __attribute__((amdgpu_waves_per_eu(2)))
k_residual_fused_batch(const BatchEnt *__restrict__ tab, int nb, Phys ph, double cm, double xk,
                       const double *__restrict__ v, long ldv, const double *__restrict__ ps,
                       const double *__restrict__ tq, const double *__restrict__ sn_all,
                       const double *__restrict__ rr, double *u, double *ff, int ityprk) {
  const unsigned wg = xcd_linear();
  const BatchEnt &E = tab[tab_find(tab, nb, BP_FUSED, wg)];
  const long np = E.b.npc;
  FusedArgs A{E.b, ph, cm, xk, v, ldv, ps, tq, sn_all + 9 * np, LIM ? rr + 15 * np : nullptr, u, ff, ityprk, E.ks};
  fused_body<NS, LIM>(A, tab_tile(E, BP_FUSED, wg));
}

// The same all-blocks residual as one persistent launch: about two
// workgroups per CU take the tiles of the flattened grid (block by block, in
// the table's order) from a queue, a device counter, until it runs dry.  The
// last workgroup to leave resets the counters, so every launch finds them at
// zero and records no memset (graph replays see the same start).  A tile's
// work is k_residual_fused_batch's for the same workgroup index, so the
// residual is the same bit for bit.  Calls that overlap on one device must
// use different workspace sets (each set has its own counters).
template <bool NS, bool LIM>
__global__ void __launch_bounds__(FT *FT) __attribute__((amdgpu_waves_per_eu(2)))
k_residual_fused_queue(const BatchEnt *__restrict__ tab, int nb, unsigned ntiles, unsigned *queue, Phys ph,
                       double cm, double xk, const double *__restrict__ v, long ldv, const double *__restrict__ ps,
                       const double *__restrict__ tq, const double *__restrict__ sn_all,
                       const double *__restrict__ rr, double *u, double *ff, int ityprk) {
  __shared__ unsigned next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const unsigned wg = next;
    __syncthreads();  // read by every thread before thread 0 takes the next tile
    if (wg >= ntiles) break;
    const BatchEnt &E = tab[tab_find(tab, nb, BP_FUSED, wg)];
    const long np = E.b.npc;
    FusedArgs A{E.b, ph, cm, xk, v, ldv, ps, tq, sn_all + 9 * np, LIM ? rr + 15 * np : nullptr, u, ff, ityprk, E.ks};
    fused_body<NS, LIM>(A, tab_tile(E, BP_FUSED, wg));
  }
  if (threadIdx.x == 0) {
    __threadfence();
    // every workgroup has left its loop once the count reaches the grid:
    // nothing takes from the queue any more
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
}

// u = 0 on the non-real entries of every block's gather range
__global__ void __launch_bounds__(256) k_zero_ghost_batch(const BatchEnt *__restrict__ tab, int nb, double *u,
                                                          long ldv) {
  const BatchEnt &E = tab[tab_find(tab, nb, BP_GATHER, blockIdx.x)];
  zero_ghost_body(E.b, blockIdx.x - E.start[BP_GATHER], E.n0, E.n1, u, ldv);
}

__global__ void __launch_bounds__(256) k_gather_batch(const BatchEnt *__restrict__ tab, int nb,
                                                      const double *__restrict__ H, double *u, long ldv, double *ff,
                                                      int ityprk) {
  const BatchEnt &E = tab[tab_find(tab, nb, BP_GATHER, blockIdx.x)];
  gather_body(E.b, blockIdx.x - E.start[BP_GATHER], E.n0, E.n1, H + 15 * E.b.npc, u, ldv, ff, ityprk);
}

extern "C" int codens_sch_roe_batch(int32_t nblk, const codens_block *blocks, const codens_phys *cp,
                                    const codens_scheme *sc, int32_t ityprk, double *u, const double *v,
                                    int64_t ldv, double *ff, const double *tq, const double *sn_all,
                                    const double *ps, void *stream) {
  const char *what = "codens_sch_roe_batch";
  if (nblk <= 0 || !blocks || !cp || !sc || !u || !v || !tq || !sn_all || !ps || (ityprk && !ff)) {
    codens_set_error("%s: null or invalid argument", what);
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  std::vector<BatchEnt> ent(nblk);
  std::vector<int32_t> key;
  unsigned tot[BP_N] = {0};
  std::vector<std::pair<long, long>> span(nblk);
  const int ilim = sc->ilim;
  for (int l = 0; l < nblk; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    const int32_t *w = (const int32_t *)&blocks[l];
    key.insert(key.end(), w, w + sizeof(codens_block) / sizeof(int32_t));
    Blk b = make_blk(&blocks[l]);
    span[l] = std::make_pair(b.npc, b.npc + b.nnn);
    BatchEnt &E = ent[l];
    E.b = b;
    E.n0 = b.idx(b.i1 - 1, b.j1 - 1, b.k1 - 1);
    E.n1 = b.idx(b.i2, b.j2, b.k2);
    long nl0 = (long)(b.j2 - b.j1) * (b.k2 - b.k1);
    long nl12 = (long)(b.i2 - b.i1) * max(b.k2 - b.k1, b.j2 - b.j1);
    unsigned g[BP_N][3] = {{(unsigned)((b.nnn + 255) / 256), 1, 1},
                           {(unsigned)((nl0 + CH_LINES - 1) / CH_LINES), 5, 1},
                           {(unsigned)((nl12 + CH_LINES - 1) / CH_LINES), 5, 2},
                           {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
                           {(unsigned)((E.n1 - E.n0 + 256) / 256), 1, 1},
                           {0, 0, 0}};
    {
      // the slab depth of the per-block launch: the block's residual is then
      // the per-block one bit for bit (a slab's first k face is formed by
      // other code than the faces it carries, which may round differently)
      dim3 fg = fused_grid(b, FS);
      E.ks = slab_depth((long)fg.x * fg.y, b.k2 - b.k1, FS);
      fg = fused_grid(b, E.ks);
      g[BP_FUSED][0] = fg.x; g[BP_FUSED][1] = fg.y; g[BP_FUSED][2] = fg.z;
    }
    for (int d = 0; d < 3; d++) {
      int nx = b.i2 - b.i1 + (d == 0), ny = b.j2 - b.j1 + (d == 1), nz = b.k2 - b.k1 + (d == 2);
      g[BP_FACE0 + d][0] = (nx + TX - 1) / TX;
      g[BP_FACE0 + d][1] = (ny + TY - 1) / TY;
      g[BP_FACE0 + d][2] = nz;
    }
    for (int p = 0; p < BP_N; p++) {
      E.start[p] = tot[p];
      E.gx[p] = g[p][0];
      E.gy[p] = g[p][1];
      tot[p] += g[p][0] * g[p][1] * g[p][2];
    }
  }
  // blocks overlapping in storage would share workspace ranges and race
  std::sort(span.begin(), span.end());
  for (int l = 0; l < nblk; l++) {
    if ((l == 0 && span[0].first != 0) || (l > 0 && span[l].first < span[l - 1].second) || span[l].second > ldv) {
      codens_set_error("%s: the blocks must start at entry 0, not overlap, and lie within ldv", what);
      return CODENS_EINVAL;
    }
  }
  BatchEnt *tab = (BatchEnt *)codens_device_table(0, key, ent.data(), sizeof(BatchEnt) * nblk, st, what);
  if (!tab) return CODENS_EHIP;
  Phys ph = make_phys(cp);
  const long tot_n = span[nblk - 1].second;
  const bool fused = g_residual_fused >= 1;
  const bool need_uvw = !fused || ilim == 1;
  double *H = fused ? nullptr : (double *)codens_workspace(sizeof(double) * 15 * tot_n, 0);
  double *R = ilim == 1 ? (double *)codens_workspace(sizeof(double) * 15 * tot_n, 1) : nullptr;
  double *uvw = need_uvw ? (double *)codens_workspace(sizeof(double) * 3 * tot_n, 4) : nullptr;
  if ((!fused && !H) || (need_uvw && !uvw) || (ilim == 1 && !R)) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  if (need_uvw) {
    CodensTimed _tm("roe_prim", st);
    hipLaunchKernelGGL(k_prim_uvw_batch, dim3(tot[BP_PRIM]), dim3(256), 0, st, (const BatchEnt *)tab, (int)nblk, v,
                       (long)ldv, uvw);
  }
  if (ilim == 1) {
    CodensTimed _tm("roe_limiter_chain", st);
    hipLaunchKernelGGL(k_chain_i_batch, dim3(tot[BP_CHAIN_I]), dim3(64), 0, st, (const BatchEnt *)tab, (int)nblk, v,
                       (const double *)uvw, ps, R);
    hipLaunchKernelGGL(k_chain_jk_batch, dim3(tot[BP_CHAIN_JK]), dim3(64), 0, st, (const BatchEnt *)tab, (int)nblk,
                       v, (const double *)uvw, ps, R);
  }
  if (fused) {
    const double cm = 0.25 * sc->muscl;
    if (g_residual_fused == 2) {
      unsigned *queue = codens_queue_counters(st, what);
      if (!queue) return CODENS_EHIP;
      const unsigned ntiles = tot[BP_FUSED];
      const unsigned nwg = std::min<unsigned>(ntiles, (unsigned)codens_resident_workgroups(2));
      CodensTimed _tm("roe_fused", st);
      if (ilim == 1)
        hipLaunchKernelGGL((k_residual_fused_queue<true, true>), dim3(nwg), dim3(FT * FT), 0, st,
                           (const BatchEnt *)tab, (int)nblk, ntiles, queue, ph, cm, sc->xk, v, (long)ldv, ps, tq,
                           sn_all, (const double *)R, u, ff, (int)ityprk);
      else
        hipLaunchKernelGGL((k_residual_fused_queue<true, false>), dim3(nwg), dim3(FT * FT), 0, st,
                           (const BatchEnt *)tab, (int)nblk, ntiles, queue, ph, cm, sc->xk, v, (long)ldv, ps, tq,
                           sn_all, (const double *)R, u, ff, (int)ityprk);
    } else {
      CodensTimed _tm("roe_fused", st);
      if (ilim == 1)
        hipLaunchKernelGGL((k_residual_fused_batch<true, true>), dim3(tot[BP_FUSED]), dim3(FT * FT), 0, st,
                           (const BatchEnt *)tab, (int)nblk, ph, cm, sc->xk, v, (long)ldv, ps, tq, sn_all,
                           (const double *)R, u, ff, (int)ityprk);
      else
        hipLaunchKernelGGL((k_residual_fused_batch<true, false>), dim3(tot[BP_FUSED]), dim3(FT * FT), 0, st,
                           (const BatchEnt *)tab, (int)nblk, ph, cm, sc->xk, v, (long)ldv, ps, tq, sn_all,
                           (const double *)R, u, ff, (int)ityprk);
    }
    CodensTimed _tm("roe_zero_ghost", st);
    hipLaunchKernelGGL(k_zero_ghost_batch, dim3(tot[BP_GATHER]), dim3(256), 0, st, (const BatchEnt *)tab, (int)nblk,
                       u, (long)ldv);
    return codens_check_launch(what);
  }
  for (int d = 0; d < 3; d++) {
    CodensTimed _tm("roe_face_flux", st);
    hipLaunchKernelGGL(k_face_flux_batch, dim3(tot[BP_FACE0 + d]), dim3(TX, TY), 0, st, (const BatchEnt *)tab,
                       (int)nblk, ph, d, 1, ilim, (double)sc->muscl, sc->xk, v, (long)ldv, ps, tq, sn_all,
                       (const double *)uvw, (const double *)R, H);
  }
  {
    CodensTimed _tm("roe_gather", st);
    hipLaunchKernelGGL(k_gather_batch, dim3(tot[BP_GATHER]), dim3(256), 0, st, (const BatchEnt *)tab, (int)nblk,
                       (const double *)H, u, (long)ldv, ff, (int)ityprk);
  }
  return codens_check_launch(what);
}
