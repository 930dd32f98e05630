// Closing part of the turbulence stage of met_num (met_num.f90:697-930) on
// gfx950: the line-implicit phase of the two transport equations
// (impli2_eqt_3d), the limiters of at_cutke (met_cut, met_cutsa) and the
// eddy viscosity of the SA and k-omega models (met_samut, met_kocmut).
#include "common.hpp"

#include <algorithm>
#include <utility>
#include <vector>

using namespace codens;

#define REAL3(b) REAL3_AT(b, xcd_block())
// the same, for the tile XB of a block's own grid
#define REAL3_AT(b, XB)                              \
  const dim3 xb = (XB);                              \
  int c[3];                                          \
  c[0] = (b).i1 + xb.x * TX + threadIdx.x;           \
  c[1] = (b).j1 + xb.y * TY + threadIdx.y;           \
  c[2] = (b).k1 + xb.z;                              \
  if (c[0] >= (b).i2 || c[1] >= (b).j2) return;

// Workspace of the implicit phase, block-local index m: rv, diag, the half
// normal velocities u1..u3 and coefficients of the faces, the elimination
// coefficients alpha, beta6, beta7 (line-major), the increments dw6, dw7, and
// the line systems of the current sweep (line-major).
struct ImpliWs {
  double *rv, *cd, *uf, *ce, *al, *b6, *b7, *dw6, *dw7;
  double *la, *lb, *lc, *l6, *l7;  // line-major coefficients and right-hand sides of one sweep
};

// the 18 planes of nnn of a block's workspace starting at ws
__host__ __device__ inline ImpliWs impli_ws(double *ws, long nnn) {
  return ImpliWs{ws,           ws + nnn,      ws + 2 * nnn,  ws + 5 * nnn,  ws + 8 * nnn,
                 ws + 9 * nnn, ws + 10 * nnn, ws + 11 * nnn, ws + 12 * nnn, ws + 13 * nnn,
                 ws + 14 * nnn, ws + 15 * nnn, ws + 16 * nnn, ws + 17 * nnn};
}

// impli2_eqt_3d.f90:113-122: rv and the diagonal's own part
__device__ __forceinline__ void cells_body(const Blk &b, dim3 tile, const double *__restrict__ v,
                                           const double *__restrict__ mu, const double *__restrict__ mut,
                                           const double *__restrict__ dt, const double *__restrict__ vol,
                                           const double *__restrict__ cfke, double cmt, const ImpliWs &w) {
  REAL3_AT(b, tile)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  w.rv[m] = (mu[n] + mut[n] / cmt) / v[n];
  w.cd[m] = vol[n] / dt[n] + cfke[n] * vol[n];
}

__global__ void __launch_bounds__(TX *TY) k_impli2_cells(Blk b, const double *__restrict__ v, long ldv,
                                                         const double *__restrict__ mu,
                                                         const double *__restrict__ mut,
                                                         const double *__restrict__ dt,
                                                         const double *__restrict__ vol,
                                                         const double *__restrict__ cfke, double cmt, ImpliWs w) {
  cells_body(b, xcd_block(), v, mu, mut, dt, vol, cfke, cmt, w);
}

// impli2_eqt_3d.f90:126-208: coefficients of every face i1..i2 (resp. j, k)
// of the real transverse range; the launch covers the cells up to i2, j2, k2.
__device__ __forceinline__ void faces_body(const Blk &b, dim3 xb, const double *__restrict__ v, long ldv,
                                           const double *__restrict__ sn, const double *__restrict__ vol,
                                           const ImpliWs &w) {
  int c[3];
  c[0] = b.i1 + xb.x * TX + threadIdx.x;
  c[1] = b.j1 + xb.y * TY + threadIdx.y;
  c[2] = b.k1 + xb.z;
  if (c[0] > b.i2 || c[1] > b.j2) return;
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    bool ok = true;
#pragma unroll
    for (int x = 0; x < 3; x++) ok = ok && c[x] <= b.hi(x) - (x == d ? 0 : 1);
    if (!ok) continue;
    long inc = b.inc(d);
    double s1 = snv(sn, b.nnn, m, d, 0), s2 = snv(sn, b.nnn, m, d, 1), s3 = snv(sn, b.nnn, m, d, 2);
    double cnds = s1 * s1 + s2 * s2 + s3 * s3;
    double uu = 0.5 * (v[ldv + n] / v[n] + v[ldv + n - inc] / v[n - inc]);
    double vv = 0.5 * (v[2 * ldv + n] / v[n] + v[2 * ldv + n - inc] / v[n - inc]);
    double ww = 0.5 * (v[3 * ldv + n] / v[n] + v[3 * ldv + n - inc] / v[n - inc]);
    double uf = 0.5 * (uu * s1 + vv * s2 + ww * s3);
    w.uf[d * b.nnn + m] = uf;
    w.ce[d * b.nnn + m] = fabs(uf) + (w.rv[m] + w.rv[m - inc]) * cnds / (vol[n] + vol[n - inc]);
  }
}

__global__ void __launch_bounds__(TX *TY) k_impli2_faces(Blk b, const double *__restrict__ v, long ldv,
                                                         const double *__restrict__ sn,
                                                         const double *__restrict__ vol, ImpliWs w) {
  faces_body(b, xcd_block(), v, ldv, sn, vol, w);
}

// impli2_eqt_3d.f90:210-220
__device__ __forceinline__ void diag_body(const Blk &b, dim3 tile, const ImpliWs &w) {
  REAL3_AT(b, tile)
  long m = b.idx(c[0], c[1], c[2]) - b.npc, nnn = b.nnn;
  const double *ce = w.ce;
  w.cd[m] = w.cd[m] + ce[m] + ce[m + 1] + ce[nnn + m] + ce[nnn + m + b.nid] + ce[2 * nnn + m] +
            ce[2 * nnn + m + b.nijd];
}

__global__ void __launch_bounds__(TX *TY) k_impli2_diag(Blk b, ImpliWs w) { diag_body(b, xcd_block(), w); }

// One sweep along direction d (impli2_eqt_3d.f90:238-275, 292-329, 346-383)
// in two launches.
//
// k_impli2_rhs, over all real cells: the line system's coefficients a, b, c
// and the right-hand sides of both equations, whose transverse terms read
// the previous sweep's increments only, so every cell is independent.  They
// go to a line-major layout, entry (a - lo_d) * nl + line with the line index
// (c_p - lo_p) + (c_q - lo_q) * n_p, so that the recursion below reads
// consecutive lines at consecutive addresses.  Tiles are 16 x 16 cells of one
// k plane; for i lines the tile is transposed through LDS, so that both the
// reads (i fastest) and the writes (j fastest) move 128-byte runs.
//
// k_impli2_line, one thread per line: the elimination from the high end of
// the line (recursion values zero beyond the last real cell, as the
// reference's zeroed arrays give) and then the substitution from the low
// end, which writes the increments dw6, dw7.  The coefficients do not depend
// on the recursion, so eight steps' loads are issued ahead of their
// arithmetic.
constexpr int RT = 16;

__device__ __forceinline__ void rhs_body(const Blk &b, dim3 xb, int d, const double *__restrict__ u, long ldv,
                                         const ImpliWs &w) {
  __shared__ double sh[5][RT][RT + 1];
  const int tx = threadIdx.x % RT, ty = threadIdx.x / RT;
  const int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
  const long np_ = b.hi(p) - b.lo(p), nl = np_ * (b.hi(q) - b.lo(q));
  int c[3];
  c[0] = b.i1 + xb.x * RT + tx;
  c[1] = b.j1 + xb.y * RT + ty;
  c[2] = b.k1 + xb.z;
  const bool in = c[0] < b.i2 && c[1] < b.j2;
  double val[5] = {0., 0., 0., 0., 0.};
  if (in) {
    const long inc = b.inc(d), ip = b.inc(p), iq = b.inc(q), nnn = b.nnn;
    const long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
    const double *__restrict__ uD = w.uf + d * nnn, *__restrict__ uP = w.uf + p * nnn;
    const double *__restrict__ uQ = w.uf + q * nnn;
    const double *__restrict__ cD = w.ce + d * nnn, *__restrict__ cP = w.ce + p * nnn;
    const double *__restrict__ cQ = w.ce + q * nnn;
    const double *__restrict__ w6 = w.dw6, *__restrict__ w7 = w.dw7;
    double td = -uP[m + ip] + uP[m] - uQ[m + iq] + uQ[m];
    double tpp = cP[m + ip] - uP[m + ip], tmp_ = cP[m] + uP[m];
    double tpq = cQ[m + iq] - uQ[m + iq], tmq = cQ[m] + uQ[m];
    val[3] = -u[5 * ldv + n] + td * w6[m] + tmp_ * w6[m - ip] + tpp * w6[m + ip] + tmq * w6[m - iq] +
             tpq * w6[m + iq];
    val[4] = -u[6 * ldv + n] + td * w7[m] + tmp_ * w7[m - ip] + tpp * w7[m + ip] + tmq * w7[m - iq] +
             tpq * w7[m + iq];
    val[0] = cD[m] + uD[m];
    val[1] = w.cd[m] + uD[m + inc] - uD[m];
    val[2] = -cD[m + inc] + uD[m + inc];
  }
  double *const out[5] = {w.la, w.lb, w.lc, w.l6, w.l7};
  if (d != 0) {
    if (in) {
      const long L = (long)(c[d] - b.lo(d)) * nl + (c[p] - b.lo(p)) + (long)(c[q] - b.lo(q)) * np_;
#pragma unroll
      for (int e = 0; e < 5; e++) out[e][L] = val[e];
    }
    return;
  }
#pragma unroll
  for (int e = 0; e < 5; e++) sh[e][ty][tx] = val[e];
  __syncthreads();
  // i lines: p = j, q = k; thread (tx, ty) writes cell (i0 + ty, j0 + tx)
  const int i = b.i1 + xb.x * RT + ty, j = b.j1 + xb.y * RT + tx;
  if (i < b.i2 && j < b.j2) {
    const long L = (long)(i - b.i1) * nl + (j - b.j1) + (long)(c[2] - b.k1) * np_;
#pragma unroll
    for (int e = 0; e < 5; e++) out[e][L] = sh[e][tx][ty];
  }
}

__global__ void __launch_bounds__(RT *RT) k_impli2_rhs(Blk b, int d, const double *__restrict__ u, long ldv,
                                                        ImpliWs w) {
  rhs_body(b, xcd_block(), d, u, ldv, w);
}

constexpr int LU = 8;  // recursion steps per batch of loads

// The two recursions of one line, each software-pipelined: the loads of the
// next LU steps are issued before the arithmetic of the current LU steps, so
// one memory round trip covers 2 x LU steps of the serial chain.
__device__ __forceinline__ void line_body(const Blk &b, int d, long t, const ImpliWs &w) {
  const int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
  const long np_ = b.hi(p) - b.lo(p), nl = np_ * (b.hi(q) - b.lo(q));
  if (t >= nl) return;
  const int nd = b.hi(d) - b.lo(d), nb = nd / LU;
  const double *A = w.la + t, *B = w.lb + t, *C = w.lc + t, *R6 = w.l6 + t, *R7 = w.l7 + t;
  double *AL = w.al + t, *B6 = w.b6 + t, *B7 = w.b7 + t;
  // elimination from the high end: steps a = nd-1, nd-2, ...
  double al1 = 0., b61 = 0., b71 = 0.;  // values at the next face up the line
  double ai[LU], bi[LU], ci[LU], r6[LU], r7[LU];
  if (nb > 0) {
#pragma unroll
    for (int x = 0; x < LU; x++) {
      const long L = (long)(nd - 1 - x) * nl;
      ai[x] = A[L]; bi[x] = B[L]; ci[x] = C[L]; r6[x] = R6[L]; r7[x] = R7[L];
    }
  }
  for (int bt = 0; bt < nb; bt++) {
    const int a = nd - 1 - bt * LU;
    double an[LU], bn[LU], cn[LU], sn6[LU], sn7[LU];
    if (bt + 1 < nb) {
#pragma unroll
      for (int x = 0; x < LU; x++) {
        const long L = (long)(a - LU - x) * nl;
        an[x] = A[L]; bn[x] = B[L]; cn[x] = C[L]; sn6[x] = R6[L]; sn7[x] = R7[L];
      }
    }
#pragma unroll
    for (int x = 0; x < LU; x++) {
      const long L = (long)(a - x) * nl;
      double den = bi[x] + ci[x] * al1;
      al1 = ai[x] / den;
      b61 = (r6[x] - ci[x] * b61) / den;
      b71 = (r7[x] - ci[x] * b71) / den;
      AL[L] = al1; B6[L] = b61; B7[L] = b71;
    }
#pragma unroll
    for (int x = 0; x < LU; x++) {
      ai[x] = an[x]; bi[x] = bn[x]; ci[x] = cn[x]; r6[x] = sn6[x]; r7[x] = sn7[x];
    }
  }
  for (int a = nd - 1 - nb * LU; a >= 0; a--) {
    const long L = (long)a * nl;
    double c0 = C[L];
    double den = B[L] + c0 * al1;
    al1 = A[L] / den;
    b61 = (R6[L] - c0 * b61) / den;
    b71 = (R7[L] - c0 * b71) / den;
    AL[L] = al1; B6[L] = b61; B7[L] = b71;
  }
  // substitution from the low end: steps a = 0, 1, ...  (same thread wrote
  // AL, B6, B7 above, so its own loads see those values)
  int c[3];
  c[p] = b.lo(p) + (int)(t % np_);
  c[q] = b.lo(q) + (int)(t / np_);
  c[d] = b.lo(d);
  const long inc = b.inc(d);
  const long m0 = b.idx(c[0], c[1], c[2]) - b.npc;
  double *dw6 = w.dw6, *dw7 = w.dw7;
  double w6 = 0., w7 = 0.;  // increments of the ghost cell below the line
  double xa[LU], x6[LU], x7[LU];
  if (nb > 0) {
#pragma unroll
    for (int x = 0; x < LU; x++) {
      const long L = (long)x * nl;
      xa[x] = AL[L]; x6[x] = B6[L]; x7[x] = B7[L];
    }
  }
  for (int bt = 0; bt < nb; bt++) {
    const int a = bt * LU;
    double ya[LU], y6[LU], y7[LU];
    if (bt + 1 < nb) {
#pragma unroll
      for (int x = 0; x < LU; x++) {
        const long L = (long)(a + LU + x) * nl;
        ya[x] = AL[L]; y6[x] = B6[L]; y7[x] = B7[L];
      }
    }
#pragma unroll
    for (int x = 0; x < LU; x++) {
      w6 = xa[x] * w6 + x6[x];
      w7 = xa[x] * w7 + x7[x];
      const long m = m0 + (long)(a + x) * inc;
      dw6[m] = w6;
      dw7[m] = w7;
    }
#pragma unroll
    for (int x = 0; x < LU; x++) {
      xa[x] = ya[x]; x6[x] = y6[x]; x7[x] = y7[x];
    }
  }
  for (int a = nb * LU; a < nd; a++) {
    const long L = (long)a * nl;
    w6 = AL[L] * w6 + B6[L];
    w7 = AL[L] * w7 + B7[L];
    dw6[m0 + (long)a * inc] = w6;
    dw7[m0 + (long)a * inc] = w7;
  }
}

__global__ void __launch_bounds__(64) k_impli2_line(Blk b, int d, ImpliWs w) {
  line_body(b, d, blockIdx.x * 64L + threadIdx.x, w);
}

// :406-416
__device__ __forceinline__ void impli2_update_body(const Blk &b, dim3 tile, double *v, long ldv, const ImpliWs &w) {
  REAL3_AT(b, tile)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  v[5 * ldv + n] = v[5 * ldv + n] + w.dw6[m];
  v[6 * ldv + n] = v[6 * ldv + n] + w.dw7[m];
}

__global__ void __launch_bounds__(TX *TY) k_impli2_update(Blk b, double *v, long ldv, ImpliWs w) {
  impli2_update_body(b, xcd_block(), v, ldv, w);
}

__global__ void k_zero_n(double *p, long cnt) {
  long i = blockIdx.x * 256L + threadIdx.x;
  if (i < cnt) p[i] = 0.;
}

extern "C" int codens_impli2_eqt_3d(const codens_block *cb, const double *u, int64_t ldv, const double *dt,
                                    double *v, const double *mu, const double *mut, const double *cfke,
                                    const double *sn, const double *vol, double cmt, void *stream) {
  int rc = codens_check_block(cb, "codens_impli2_eqt_3d");
  if (rc) return rc;
  if (!u || !dt || !v || !mu || !mut || !cfke || !sn || !vol || !(cmt > 0.)) {
    codens_set_error("codens_impli2_eqt_3d: invalid argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipStream_t st = (hipStream_t)stream;
  long nnn = b.nnn;
  double *ws = (double *)codens_workspace(sizeof(double) * 18 * nnn, 3);
  if (!ws) {
    codens_set_error("codens_impli2_eqt_3d: workspace allocation failed");
    return CODENS_EHIP;
  }
  ImpliWs w = impli_ws(ws, nnn);
  CodensTimed _tm("impli2_eqt_3d", st);
  // rv, the face arrays and the increments start at zero (ghost entries stay)
  hipLaunchKernelGGL(k_zero_n, dim3((unsigned)((13 * nnn + 255) / 256)), dim3(256), 0, st, ws, 13 * nnn);
  dim3 g = cell_grid(b), t(TX, TY);
  hipLaunchKernelGGL(k_impli2_cells, g, t, 0, st, b, v, (long)ldv, mu, mut, dt, vol, cfke, cmt, w);
  dim3 gf((b.i2 - b.i1 + 1 + TX - 1) / TX, (b.j2 - b.j1 + 1 + TY - 1) / TY, b.k2 - b.k1 + 1);
  hipLaunchKernelGGL(k_impli2_faces, gf, t, 0, st, b, v, (long)ldv, sn, vol, w);
  hipLaunchKernelGGL(k_impli2_diag, g, t, 0, st, b, w);
  dim3 gr((b.i2 - b.i1 + RT - 1) / RT, (b.j2 - b.j1 + RT - 1) / RT, b.k2 - b.k1);
  for (int d = 0; d < 3; d++) {
    int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
    int nl = (b.hi(p) - b.lo(p)) * (b.hi(q) - b.lo(q));
    for (int sweep = 0; sweep < 2; sweep++) {
      hipLaunchKernelGGL(k_impli2_rhs, gr, dim3(RT * RT), 0, st, b, d, u, (long)ldv, w);
      hipLaunchKernelGGL(k_impli2_line, dim3((nl + 63) / 64), dim3(64), 0, st, b, d, w);
    }
  }
  hipLaunchKernelGGL(k_impli2_update, g, t, 0, st, b, v, (long)ldv, w);
  return codens_check_launch("codens_impli2_eqt_3d");
}

// ---------------------------------------------------------------------------
// impli2_eqt_3d for all blocks of a layout: each pass of the phase (zero,
// cells, faces, diagonal, and per sweep the line systems and the line
// recursions, then the update) is one flattened launch over every block,
// with a device table of the blocks' extents and first workgroups per pass.
// Block l's 18-plane workspace starts at 18 npc.  The line recursion is a
// serial chain per line, so a launch takes about the time of its longest
// line: one launch over the lines of all blocks replaces one per block.
enum { TB_ZERO, TB_CELL, TB_FACE, TB_RHS, TB_LINE0, TB_LINE1, TB_LINE2, TB_N };

struct TurbEnt {
  Blk b;
  unsigned start[TB_N];         // first workgroup of this block, per pass
  unsigned gx[TB_N], gy[TB_N];  // this block's grid, per pass
};

__global__ void __launch_bounds__(256) k_impli2_zero_batch(const TurbEnt *__restrict__ tab, int nb, double *ws) {
  const TurbEnt &E = tab[tab_find(tab, nb, TB_ZERO, blockIdx.x)];
  const long i = (blockIdx.x - E.start[TB_ZERO]) * 256L + threadIdx.x;
  if (i < 13 * E.b.nnn) ws[18 * E.b.npc + i] = 0.;
}

__global__ void __launch_bounds__(TX *TY) k_impli2_cells_batch(const TurbEnt *__restrict__ tab, int nb,
                                                               const double *__restrict__ v,
                                                               const double *__restrict__ mu,
                                                               const double *__restrict__ mut,
                                                               const double *__restrict__ dt,
                                                               const double *__restrict__ vol,
                                                               const double *__restrict__ cfke, double cmt,
                                                               double *ws) {
  const unsigned wg = xcd_linear();
  const TurbEnt &E = tab[tab_find(tab, nb, TB_CELL, wg)];
  const Blk b = E.b;
  cells_body(b, tab_tile(E, TB_CELL, wg), v, mu, mut, dt, vol, cfke, cmt, impli_ws(ws + 18 * b.npc, b.nnn));
}

__global__ void __launch_bounds__(TX *TY) k_impli2_faces_batch(const TurbEnt *__restrict__ tab, int nb,
                                                               const double *__restrict__ v, long ldv,
                                                               const double *__restrict__ sn_all,
                                                               const double *__restrict__ vol, double *ws) {
  const unsigned wg = xcd_linear();
  const TurbEnt &E = tab[tab_find(tab, nb, TB_FACE, wg)];
  const Blk b = E.b;
  faces_body(b, tab_tile(E, TB_FACE, wg), v, ldv, sn_all + 9 * b.npc, vol, impli_ws(ws + 18 * b.npc, b.nnn));
}

__global__ void __launch_bounds__(TX *TY) k_impli2_diag_batch(const TurbEnt *__restrict__ tab, int nb, double *ws) {
  const unsigned wg = xcd_linear();
  const TurbEnt &E = tab[tab_find(tab, nb, TB_CELL, wg)];
  const Blk b = E.b;
  diag_body(b, tab_tile(E, TB_CELL, wg), impli_ws(ws + 18 * b.npc, b.nnn));
}

__global__ void __launch_bounds__(RT *RT) k_impli2_rhs_batch(const TurbEnt *__restrict__ tab, int nb, int d,
                                                             const double *__restrict__ u, long ldv, double *ws) {
  const unsigned wg = xcd_linear();
  const TurbEnt &E = tab[tab_find(tab, nb, TB_RHS, wg)];
  const Blk b = E.b;
  rhs_body(b, tab_tile(E, TB_RHS, wg), d, u, ldv, impli_ws(ws + 18 * b.npc, b.nnn));
}

__global__ void __launch_bounds__(64) k_impli2_line_batch(const TurbEnt *__restrict__ tab, int nb, int d,
                                                          double *ws) {
  const TurbEnt &E = tab[tab_find(tab, nb, TB_LINE0 + d, blockIdx.x)];
  const Blk b = E.b;
  line_body(b, d, (blockIdx.x - E.start[TB_LINE0 + d]) * 64L + threadIdx.x, impli_ws(ws + 18 * b.npc, b.nnn));
}

__global__ void __launch_bounds__(TX *TY) k_impli2_update_batch(const TurbEnt *__restrict__ tab, int nb, double *v,
                                                                long ldv, double *ws) {
  const unsigned wg = xcd_linear();
  const TurbEnt &E = tab[tab_find(tab, nb, TB_CELL, wg)];
  const Blk b = E.b;
  impli2_update_body(b, tab_tile(E, TB_CELL, wg), v, ldv, impli_ws(ws + 18 * b.npc, b.nnn));
}

extern "C" int codens_impli2_eqt_3d_batch(int32_t nblk, const codens_block *blocks, const double *u, int64_t ldv,
                                          const double *dt, double *v, const double *mu, const double *mut,
                                          const double *cfke, const double *sn_all, const double *vol, double cmt,
                                          void *stream) {
  const char *what = "codens_impli2_eqt_3d_batch";
  if (nblk <= 0 || !blocks || !u || !dt || !v || !mu || !mut || !cfke || !sn_all || !vol || !(cmt > 0.)) {
    codens_set_error("%s: null or invalid argument", what);
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  std::vector<TurbEnt> ent(nblk);
  std::vector<int32_t> key;
  std::vector<std::pair<long, long>> span(nblk);
  unsigned tot[TB_N] = {0};
  for (int l = 0; l < nblk; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    const int32_t *wd = (const int32_t *)&blocks[l];
    key.insert(key.end(), wd, wd + sizeof(codens_block) / sizeof(int32_t));
    Blk b = make_blk(&blocks[l]);
    span[l] = {b.npc, b.npc + b.nnn};
    TurbEnt &E = ent[l];
    E.b = b;
    const unsigned ni = b.i2 - b.i1, nj = b.j2 - b.j1, nk = b.k2 - b.k1;
    dim3 cg = cell_grid(b);
    unsigned g[TB_N][3] = {{(unsigned)((13 * b.nnn + 255) / 256), 1, 1},
                           {cg.x, cg.y, cg.z},
                           {(ni + 1 + TX - 1) / TX, (nj + 1 + TY - 1) / TY, nk + 1},
                           {(ni + RT - 1) / RT, (nj + RT - 1) / RT, nk},
                           {(nj * nk + 63) / 64, 1, 1},
                           {(ni * nk + 63) / 64, 1, 1},
                           {(ni * nj + 63) / 64, 1, 1}};
    for (int p = 0; p < TB_N; p++) {
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
  TurbEnt *tab = (TurbEnt *)codens_device_table(2, key, ent.data(), sizeof(TurbEnt) * nblk, st, what);
  if (!tab) return CODENS_EHIP;
  double *ws = (double *)codens_workspace(sizeof(double) * 18 * tot_n, 3);
  if (!ws) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  const TurbEnt *ct = tab;
  const int nb = nblk;
  CodensTimed _tm("impli2_eqt_3d", st);
  hipLaunchKernelGGL(k_impli2_zero_batch, dim3(tot[TB_ZERO]), dim3(256), 0, st, ct, nb, ws);
  dim3 t(TX, TY);
  hipLaunchKernelGGL(k_impli2_cells_batch, dim3(tot[TB_CELL]), t, 0, st, ct, nb, (const double *)v, mu, mut, dt, vol,
                     cfke, cmt, ws);
  hipLaunchKernelGGL(k_impli2_faces_batch, dim3(tot[TB_FACE]), t, 0, st, ct, nb, (const double *)v, (long)ldv,
                     sn_all, vol, ws);
  hipLaunchKernelGGL(k_impli2_diag_batch, dim3(tot[TB_CELL]), t, 0, st, ct, nb, ws);
  for (int d = 0; d < 3; d++) {
    for (int sweep = 0; sweep < 2; sweep++) {
      hipLaunchKernelGGL(k_impli2_rhs_batch, dim3(tot[TB_RHS]), dim3(RT * RT), 0, st, ct, nb, d, u, (long)ldv, ws);
      hipLaunchKernelGGL(k_impli2_line_batch, dim3(tot[TB_LINE0 + d]), dim3(64), 0, st, ct, nb, d, ws);
    }
  }
  hipLaunchKernelGGL(k_impli2_update_batch, dim3(tot[TB_CELL]), t, 0, st, ct, nb, v, (long)ldv, ws);
  return codens_check_launch(what);
}

// met_samut.f90:46-58
__global__ void __launch_bounds__(TX *TY) k_met_samut(Blk b, double cv13, const double *__restrict__ v, long ldv,
                                                      const double *__restrict__ mu, double *__restrict__ mut) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]);
  double x = v[5 * ldv + n] / mu[n];
  double x3 = x * x * x;
  double fv1 = x3 / (x3 + cv13);
  mut[n] = v[5 * ldv + n] * fv1;
}

// met_kocmut.f90:53-90
__global__ void __launch_bounds__(TX *TY) k_met_kocmut(Blk b, int kwsst, double betas, const double *__restrict__ v,
                                                       long ldv, const double *__restrict__ mu,
                                                       double *__restrict__ mut, const double *__restrict__ dist,
                                                       const double *__restrict__ g9) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc, nnn = b.nnn;
  if (kwsst == 0) {
    mut[n] = v[n] * v[5 * ldv + n] / v[6 * ldv + n];
    return;
  }
  double a1 = sqrt(betas);
  double omeg = v[6 * ldv + n] / v[n];
  double coef1 = sqrt(v[5 * ldv + n] / v[n]) / (betas * omeg * dist[n]);
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

// met_cutsa.f90:40-49 (which = 2) and met_cut.f90:41-53 (which = 1)
__global__ void __launch_bounds__(TX *TY) k_met_cut(Blk b, int which, double epsk, double epse, double *v, long ldv) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]);
  double *s6 = v + 5 * ldv, *s7 = v + 6 * ldv;
  if (which == 2) {
    if (s6[n] < 0.) s6[n] = epsk;
    s7[n] = 0.;
  } else {
    double t6 = copysign(1.0, s6[n] - epsk), t7 = copysign(1.0, s7[n] - epse);
    double ts = copysign(1.0, t6 + t7 - 1.5);
    s6[n] = 0.5 * ((1. + ts) * s6[n] + (1. - ts) * epsk);
    s7[n] = 0.5 * ((1. + ts) * s7[n] + (1. - ts) * epse);
  }
}

extern "C" int codens_teq_gradv(const codens_block *cb, const double *sn, const double *vol, double *v, int64_t ldv,
                                double *g9, void *stream);

extern "C" int codens_met_samut(const codens_block *cb, const double *tc, const double *v, int64_t ldv,
                                const double *mu, double *mut, void *stream) {
  int rc = codens_check_block(cb, "codens_met_samut");
  if (rc) return rc;
  if (!tc || !v || !mu || !mut) {
    codens_set_error("codens_met_samut: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  double cv1 = tc[7];
  hipLaunchKernelGGL(k_met_samut, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, cv1 * cv1 * cv1, v,
                     (long)ldv, mu, mut);
  return codens_check_launch("codens_met_samut");
}

extern "C" int codens_met_kocmut(const codens_block *cb, const double *tc, int32_t kwsst, const double *sn,
                                 const double *vol, double *v, int64_t ldv, const double *mu, double *mut,
                                 const double *dist, double *g9, void *stream) {
  int rc = codens_check_block(cb, "codens_met_kocmut");
  if (rc) return rc;
  if (!tc || !v || !mu || !mut || (kwsst == 1 && (!sn || !vol || !dist || !g9)) || (kwsst != 0 && kwsst != 1)) {
    codens_set_error("codens_met_kocmut: invalid argument");
    return CODENS_EINVAL;
  }
  if (kwsst == 1) {  // the reference computes the velocity gradients here (met_kocmut.f90:55-61)
    rc = codens_teq_gradv(cb, sn, vol, v, ldv, g9, stream);
    if (rc) return rc;
  }
  Blk b = make_blk(cb);
  hipLaunchKernelGGL(k_met_kocmut, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, (int)kwsst, tc[10], v,
                     (long)ldv, mu, mut, dist, g9);
  return codens_check_launch("codens_met_kocmut");
}

extern "C" int codens_met_cut(const codens_block *cb, int32_t which, double epsk, double epse, double *v,
                              int64_t ldv, void *stream) {
  int rc = codens_check_block(cb, "codens_met_cut");
  if (rc) return rc;
  if (!v || (which != 1 && which != 2)) {
    codens_set_error("codens_met_cut: invalid argument (which = 1 met_cut, 2 met_cutsa)");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipLaunchKernelGGL(k_met_cut, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, (int)which, epsk, epse, v,
                     (long)ldv);
  return codens_check_launch("codens_met_cut");
}

// met_smkor.f90:47-70: Wilcox k-omega sources with cross diffusion and the
// strain-rate limiter; fd = gradients of k and omega, g9 = velocity gradients
// and tprod block-local.
__global__ void __launch_bounds__(TX *TY) k_met_smkor(Blk b, double beta, double betas, double sigmw, double sigmd,
                                                      const double *__restrict__ v, long ldv,
                                                      const double *__restrict__ fd,
                                                      const double *__restrict__ g9,
                                                      const double *__restrict__ tprod, double *__restrict__ cfke,
                                                      double *__restrict__ ts6, double *__restrict__ ts7) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc, nnn = b.nnn;
  double xw = beta / betas - sigmw * (0.41 * 0.41) / sqrt(betas);
  double rcmu = 1. / sqrt(0.09);
  double gkgo = fd[3 * ldv + n] * fd[n] + fd[4 * ldv + n] * fd[ldv + n] + fd[5 * ldv + n] * fd[2 * ldv + n];
  double omeg = v[6 * ldv + n] / v[n];
  double cd = sigmd * v[n] * (gkgo > 0. ? gkgo : 0.) / omeg;
  double dxx = g9[m], dyy = g9[4 * nnn + m], dzz = g9[8 * nnn + m];
  double a = g9[7 * nnn + m] + g9[5 * nnn + m], bb = g9[2 * nnn + m] + g9[6 * nnn + m];
  double cc = g9[3 * nnn + m] + g9[nnn + m];
  double ss = sqrt(4. * (dxx * dxx + dyy * dyy + dzz * dzz) / 3. + a * a + bb * bb + cc * cc) / omeg;
  double r = rcmu / ss;
  double alpha = 1. < r ? 1. : r;
  ts6[n] = tprod[m] - betas * v[5 * ldv + n] * omeg;
  ts7[n] = xw * tprod[m] * v[6 * ldv + n] / v[5 * ldv + n] / alpha - beta * v[6 * ldv + n] * omeg + cd;
  cfke[n] = 0.18 * omeg;
}

// met_smch.f90:88-108: Chien low-Reynolds k-epsilon sources
__global__ void __launch_bounds__(TX *TY) k_met_smch(Blk b, double cke1, double cke2, double utaumin,
                                                     const double *__restrict__ v, long ldv,
                                                     const double *__restrict__ mu,
                                                     const double *__restrict__ dist,
                                                     const int32_t *__restrict__ mnpar,
                                                     const double *__restrict__ utau,
                                                     const double *__restrict__ tprod, double *__restrict__ ts6,
                                                     double *__restrict__ ts7) {
  REAL3(b)
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  int mp = mnpar[n];
  double yplus = 1000.;
  if (mp >= 1) {
    double ut = fabs(utau[mp - 1]);
    yplus = (ut > utaumin ? ut : utaumin) * dist[n] * v[n] / mu[n];
  }
  double rodst = v[n] * (dist[n] * dist[n]);
  double k = v[5 * ldv + n], e = v[6 * ldv + n];
  ts6[n] = tprod[m] - e - 2. * mu[n] * k / rodst;
  double epssk = e / k;
  double rtur = k / (mu[n] * epssk);
  double f2 = 1. - .22222222 * exp(-rtur * rtur / 36.);
  ts7[n] = cke1 * epssk * tprod[m] - cke2 * f2 * epssk * e - 2. * mu[n] * (e / rodst) * exp(-yplus / 2.);
}

extern "C" int codens_met_smkor(const codens_block *cb, const double *tc, double sigmd, const double *v,
                                int64_t ldv, const double *fd, const double *g9, const double *tprod, double *cfke,
                                double *ts6, double *ts7, void *stream) {
  int rc = codens_check_block(cb, "codens_met_smkor");
  if (rc) return rc;
  if (!tc || !v || !fd || !g9 || !tprod || !cfke || !ts6 || !ts7) {
    codens_set_error("codens_met_smkor: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipLaunchKernelGGL(k_met_smkor, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, tc[11], tc[10], tc[13],
                     sigmd, v, (long)ldv, fd, g9, tprod, cfke, ts6, ts7);
  return codens_check_launch("codens_met_smkor");
}

extern "C" int codens_met_smch(const codens_block *cb, double cke1, double cke2, double utaumin, const double *v,
                               int64_t ldv, const double *mu, const double *dist, const int32_t *mnpar,
                               const double *utau, const double *tprod, double *ts6, double *ts7, void *stream) {
  int rc = codens_check_block(cb, "codens_met_smch");
  if (rc) return rc;
  if (!v || !mu || !dist || !mnpar || !utau || !tprod || !ts6 || !ts7) {
    codens_set_error("codens_met_smch: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  hipLaunchKernelGGL(k_met_smch, cell_grid(b), dim3(TX, TY), 0, (hipStream_t)stream, b, cke1, cke2, utaumin, v,
                     (long)ldv, mu, dist, mnpar, utau, tprod, ts6, ts7);
  return codens_check_launch("codens_met_smch");
}
