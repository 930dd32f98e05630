/* CPU restatement of the CodeNS residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline
 * leg of bench.py use it as the checker.  The product never links it.
 *
 * Arrays use the reference storage (crdms.f90:126-153): SoA over all blocks,
 * leading dimension ldv (= ip11) for multi-variable arrays, 0-based flat
 * index npc + (i-id1) + (j-jd1)*nid + (k-kd1)*nid*njd.  sn points at the
 * block's own metric slice sn(lgsnlt,3,3) (atsch_num.f90:623).
 */
#ifndef CODENS_ORACLE_H
#define CODENS_ORACLE_H

/* A loop over the faces of direction d (index ranges lo..hi) whose body
 * writes the cells on both sides of each face, or runs a recurrence along
 * d: the OpenMP threads share a direction other than d (k for d = 0, 1; j
 * for d = 2), so no two threads touch one cell or one grid line, and each
 * cell still receives its faces in increasing order along d, as in the
 * serial loop: the result is the serial one bit for bit, for any thread
 * count.  Declares i, j, k for the body. */
/* calloc(n, 8) with the zeroing shared among the OpenMP threads (the pages
 * are first touched by the threads that use them) */
#include <stdlib.h>
static inline double *orc_zeros(long n) {
  double *p = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
#pragma omp parallel for schedule(static)
  for (long q = 0; q < n; q++) p[q] = 0.;
  return p;
}

#define FOR3D(lo, hi, d) \
  _Pragma("omp parallel for schedule(static)") \
  for (int o_ = (d) == 2 ? (lo)[1] : (lo)[2]; o_ <= ((d) == 2 ? (hi)[1] : (hi)[2]); o_++) \
    for (int p_ = (d) == 2 ? (lo)[2] : (lo)[1]; p_ <= ((d) == 2 ? (hi)[2] : (hi)[1]); p_++) \
      for (int i = (lo)[0], j = (d) == 2 ? o_ : p_, k = (d) == 2 ? p_ : o_; i <= (hi)[0]; i++)
#include <stdint.h>

typedef struct {
  int32_t npc, id1, id2, jd1, jd2, kd1, kd2;
  int32_t ii1, ii2, jj1, jj2, kk1, kk2;
  int32_t nnn, pad0, pad1;
} orc_block;

typedef struct {
  double gam, gam1, pinfl, pr, prt, cp, reynz, tnz;
  int32_t iflu, pad;
} orc_phys;

typedef struct {
  int32_t muscl, ilim;
  double xk;
} orc_sch;

/* Turbulence model constants, module `modeleturb` (modules.f90:168-190),
 * set per model by atlecdon.f90:112-400. */
typedef struct {
  double cb1, cb2, sigma, kappa, cw1, cw2, cw3, cv1, ct3, ct4; /* SA */
  double betas, beta, sigmk, sigmw;                            /* Wilcox k-omega */
  double alfak, alfae;                                         /* k-eps / k-l */
} orc_turb;

/* Menter k-omega constants (atlecdon.f90:255-278): inner (1) and outer (2)
 * sets blended with the function frac. */
typedef struct {
  double betae, okappa, sigme1, sigma1, beta1, wsig1, sigme2, sigma2, beta2, wsig2;
} orc_menter;

/* One wall-normal line of met_mtcorf1 (the quantities atindnor.f90:4 gives
 * for a wall point): first cell nmin (0-based storage index), m3 range and
 * direction, cell stride dm3, wall-point window (mpn0, mpn1]. */
typedef struct {
  int64_t nmin;
  int32_t m3min, m3max, isens3, dm3, mpn0, mpn1;
} orc_wall_line;

#endif
