/* codens.h -- C-ABI of the MI355X (gfx950) residual-and-update path of CodeNS.
 *
 * Every entry point replaces one reference Fortran routine (cited on each
 * declaration).  Arrays are device pointers in the reference storage:
 *
 *  - one flat storage over all blocks, i fastest, ghost layers included
 *    (crdms.f90:126-153); a block's entry n is npc + (i-id1) + (j-jd1)*nid
 *    + (k-kd1)*nid*njd, 0-based;
 *  - multi-variable arrays are SoA with leading dimension ldv = ip11
 *    (v(ip11,ip60), solve.f90:982): variable e of entry n is at e*ldv + n;
 *  - sn_blk is the block's own metric slice sn(lgsnlt,3,3) (snorm.f90:65-71),
 *    i.e. sn + 9*npc of the global array (atsch_num.f90:623); component c of
 *    the face family d (0=i, 1=j, 2=k) of entry m is sn_blk[(c*3+d)*nnn + m];
 *  - tq is the 9-array bundle toxx,toxy,toxz,toyy,toyz,tozz,qcx,qcy,qcz of
 *    leading dimension ldv (zfluto.f90:104-121).
 *
 * All calls are asynchronous on `stream` (a hipStream_t; NULL = default
 * stream) unless stated otherwise.  Return value: 0 on success, otherwise a
 * CODENS_E* code; codens_last_error() gives the message.
 */
#ifndef CODENS_H
#define CODENS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Extents of one block, the `maillage` module entries for block lm
 * (modules.f90:31-44, set by crdms.f90:126-153). */
typedef struct codens_block {
  int32_t npc;                      /* offset of the block's first entry (= npn) */
  int32_t id1, id2, jd1, jd2, kd1, kd2; /* storage extents, ghosts included */
  int32_t ii1, ii2, jj1, jj2, kk1, kk2; /* node extents; real cells ii1..ii2-1 */
  int32_t nnn;                      /* entries of the block (= lgsnlt) */
  int32_t pad0, pad1;
} codens_block;

/* Gas constants, module `proprieteflu` (modules.f90:116-122) and tnz
 * (module `definition`).  cp = 1/(gam-1) (dfnzst.f90:67). */
typedef struct codens_phys {
  double gam, gam1, pinfl, pr, prt, cp, reynz, tnz;
  int32_t iflu, pad;
} codens_phys;

/* MUSCL switches, module `schemanum` (modules.f90:151-166). */
typedef struct codens_scheme {
  int32_t muscl, ilim;
  double xk;
} codens_scheme;

/* One wall-normal line of met_mtcorf1, as atindnor.f90:4 describes a wall
 * point: first cell nmin (0-based storage index), the along-line index m3
 * from m3min to m3max in direction isens3 (+1 / -1), cell stride dm3 (the
 * reference's cells are nmin + (m3 - m3min) * isens3 * dm3), and the
 * window (mpn0, mpn1] of the wall's point numbers in mnpar. */
typedef struct codens_wall_line {
  int64_t nmin;
  int32_t m3min, m3max, isens3, dm3, mpn0, mpn1;
} codens_wall_line;

enum {
  CODENS_OK = 0,
  CODENS_EINVAL = 1,     /* bad extents, null pointer, size over a limit */
  CODENS_EHIP = 2,       /* a HIP call or launch failed */
  CODENS_ENONPHYS = 3,   /* non-physical state (T <= 0), where the reference aborts */
  CODENS_ESTOP = 4,      /* global time step below dt1min, where the reference stops */
  CODENS_EIO = 5         /* a file could not be opened, written or read as expected */
};

const char *codens_last_error(void);

/* zpres.f90:4 -- p, T, c of the real cells. */
int codens_zpres(const codens_block *b, const codens_phys *ph, const double *v, int64_t ldv,
                 double *ps, double *temp, double *cson, void *stream);

/* zvismo.f90:4 -- molecular viscosity of the real cells (Sutherland for
 * iflu=1).  Counts cells with T <= 0 into *nbad_dev (device int, may be NULL);
 * the reference aborts on any (zvismo.f90:122-127). */
int codens_zvismo(const codens_block *b, const codens_phys *ph, const double *temp, double *mu,
                  int32_t *nbad_dev, void *stream);

/* chronos.f90:4 -- local time step of the real cells; ns != 0 adds the
 * viscous limit (equat 'ns', chronos.f90:177-193). */
int codens_chronos(const codens_block *b, const codens_phys *ph, double eta, int32_t ns,
                   const double *v, int64_t ldv, const double *mu, const double *mut,
                   const double *sn_blk, const double *vol, const double *cson, double *dt,
                   void *stream);

/* chrono.f90:105-153 (kdtl=0) -- global step over nb blocks.  Synchronous:
 * *dtmin_out receives the smallest local step; returns CODENS_ESTOP (dt left
 * untouched) when it is below dt1min, else sets every stored entry of every
 * block to dt1min. */
int codens_chrono_global(const codens_block *blocks, int32_t nb, double dt1min, double *dt,
                         double *dtmin_out, void *stream);

/* sch_roe_euler.f90:4 -- Euler Roe residual u = sum of outgoing F.S with
 * unlimited kappa-MUSCL; ityprk != 0 also does ff -= u (:976-990). */
int codens_sch_roe_euler(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                         int32_t ityprk, double *u, const double *v, int64_t ldv, double *ff,
                         const double *sn_blk, const double *ps, void *stream);

/* sch_roe.f90:4 -- Navier-Stokes Roe residual: MUSCL (minmod with the
 * chained slope ratios when ilim=1) and face-averaged stresses tq. */
int codens_sch_roe(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                   int32_t ityprk, double *u, const double *v, int64_t ldv, double *ff,
                   const double *tq, const double *sn_blk, const double *ps, void *stream);

/* zvisqc.f90:134-178 (zgrad2.f90:4 then zfluto.f90:4) -- stresses and heat
 * flux tq of the real cells from Green-Gauss gradients of u, v, w, T. */
int codens_zvisqc(const codens_block *b, const codens_phys *ph, const double *sn_blk,
                  const double *vol, const double *v, int64_t ldv, const double *temp,
                  const double *mu, const double *mut, double *tq, void *stream);

/* dissip_jameson.f90:4 -- JST artificial dissipation d(1:5) of the central
 * schemes: pressure sensor per direction, eps2 = ki2*max(nu), eps4 =
 * max(0, ki4-eps2), spectral radius |V.S|+c|S|, first-order boundary faces.
 * ki2, ki4 are the block's coefficients (ki2(lm), ki4(lm)). */
int codens_dissip_jameson(const codens_block *b, double ki2, double ki4, const double *v, int64_t ldv,
                          double *d, const double *sn_blk, const double *ps, const double *cson,
                          void *stream);

/* met_roe2o.f90:4 (klroe=2) -- second-order upwind dissipation of the two
 * transported turbulence variables t(:,6:7) into d(:,6:7): per face
 * 0.5|V.S| (delta - minmod of three face differences), differences mirrored
 * one face beyond the block.  t, d: 7 arrays of ldv. */
int codens_met_roe2o(const codens_block *b, const double *t, int64_t ldv, double *d, const double *sn_blk,
                     void *stream);

/* met_roe.f90:4 (klroe=1) -- first-order upwind dissipation |V.n| dt of
 * t(:,6:7), with the reference's index function that lands one storage entry
 * before the intended cell (met_roe.f90:184-188); the entry just before the
 * block, which the reference also zeroes, is left alone. */
int codens_met_roe(const codens_block *b, const double *t, int64_t ldv, double *d, const double *sn_blk,
                   void *stream);

/* Turbulence chain of met_num.f90:162-567.  tc points at the 16 model
 * constants of module `modeleturb` (atlecdon.f90:112-400) in the order
 * cb1 cb2 sigma kappa cw1 cw2 cw3 cv1 ct3 ct4 betas beta sigmk sigmw alfak
 * alfae.  g9 (velocity gradients dvxx..dvzz) and tprod are block-local arrays
 * of nnn entries like the reference's ip00 scratch; fd is the 6-array bundle
 * fd5x fd5y fd5z fd6x fd6y fd6z of leading dimension ldv. */

/* teq_gradv.f90:4 -- clamps rho >= 1e-20 on the block box (a side effect
 * the reference has, :79-86), then Green-Gauss gradients of u, v, w. */
int codens_teq_gradv(const codens_block *b, const double *sn_blk, const double *vol, double *v, int64_t ldv,
                     double *g9, void *stream);
/* met_prod.f90:4 -- production of k, tau^R : grad u. */
int codens_met_prod(const codens_block *b, const double *mut, const double *v, int64_t ldv, const double *g9,
                    double *tprod, void *stream);
/* met_gradtr.f90:4 -- fd = grad(v6/rho), grad(v7/rho). */
int codens_met_gradtr(const codens_block *b, const double *sn_blk, const double *vol, const double *v,
                      int64_t ldv, double *fd, void *stream);
/* met_smsa.f90:4 with met_difsa.f90:4 -- Spalart-Allmaras source ts6 (ts7=0)
 * and implicit coefficient cfke; fd holds grad(nu~) from met_gradtr. */
int codens_met_smsa(const codens_block *b, const double *tc, const double *sn_blk, const double *vol,
                    const double *v, int64_t ldv, const double *mu, const double *dist, const double *g9,
                    const double *fd, double *cfke, double *ts6, double *ts7, void *stream);
/* met_smdes.f90:4 -- the Spalart-Allmaras sources of codens_met_smsa with
 * the DES length min(d, 0.65 delta) for the wall distance and |fv2|; delta
 * (block-local, nnn, written) is met_dist.f90:66-103's longest cell edge
 * from the node coordinates x, y, z (stored like the cells, ldv). */
int codens_met_smdes(const codens_block *b, const double *tc, const double *x, const double *y, const double *z,
                     double *delta, const double *sn, const double *vol, const double *v, int64_t ldv,
                     const double *mu, const double *dist, const double *g9, const double *fd, double *cfke,
                     double *ts6, double *ts7, void *stream);

/* met_smsasas.f90:4 -- Spalart-Allmaras with the SAS destruction: the
 * sources of codens_met_smsa without the rtilde clamp and with |fv2|, and
 * where 1 - tanh((8 rtilde)^3) > 0.8 the destruction from the von Karman
 * length kappa S / sqrt(tn1^2 + tn2^2), tn1, tn2 the Laplacians of u and
 * v that met_laplaciens.f90:4 forms from grad u and grad v (ischema = 1:
 * its plain form, else the weighted form; the weights are 1). */
int codens_met_smsasas(const codens_block *b, const double *tc, int32_t ischema, const double *sn, const double *vol,
                       const double *v, int64_t ldv, const double *mu, const double *dist, const double *g9,
                       const double *fd, double *cfke, double *ts6, double *ts7, void *stream);

/* ---- Smith k-l model: kc = cklb1, ckle2, sigmal, xkappa (module
 * modeleturb, atlecdon.f90:226-231) ---- */

/* met_smkl.f90:4 (with met_bark.f90:4) -- k-l sources ts6, ts7 and the
 * implicit coefficient cfke (the larger modulus of the source Jacobian's
 * eigenvalues); fd: gradients of rho k and rho l (6 planes of ldv), g9 and
 * tprod block-local. */
int codens_met_smkl(const codens_block *b, const double *kc, const double *sn, const double *vol, const double *v,
                    int64_t ldv, const double *mu, const double *mut, const double *dist, const double *fd,
                    const double *g9, const double *tprod, double *cfke, double *ts6, double *ts7, void *stream);

/* met_klmut.f90:4 -- k-l eddy viscosity. */
int codens_met_klmut(const codens_block *b, const double *kc, const double *v, int64_t ldv, const double *mu,
                     double *mut, const double *dist, void *stream);

/* met_smko.f90:4 -- compressible Wilcox k-omega sources. */
int codens_met_smko(const codens_block *b, const double *tc, const double *v, int64_t ldv, const double *tprod,
                    double *cfke, double *ts6, double *ts7, const double *cson, void *stream);
/* met_fludc.f90:4 (kfludis=1), met_fludcsa.f90:4 (2), met_fludko.f90:4 (5)
 * -- gradients fd become diffusive flux densities, in place.  v is read for
 * kfludis = 2 only, mut for 1 and 5 only (either may be NULL otherwise). */
int codens_met_flud(const codens_block *b, int32_t kfludis, const double *tc, const double *v, int64_t ldv,
                    const double *mu, const double *mut, double *fd, void *stream);
/* sch_turb.f90:4 -- flux balance u6, u7 = 0.5 sum of centred face fluxes
 * - vol ts - d; the wall-layer zeroing met_pardis (kparoi=1) is applied by
 * the caller to d beforehand. */
int codens_sch_turb(const codens_block *b, double *u, const double *v, int64_t ldv, const double *d,
                    const double *fd, const double *ts6, const double *ts7, const double *sn_blk,
                    const double *vol, void *stream);

/* impli2_eqt_3d.f90:4 -- implicit phase of the two transport equations:
 * face coefficients |0.5 V.S| + (rv_L + rv_R)|S|^2/(vol_L + vol_R) with
 * rv = (mu + mut/cmt)/rho, diagonal vol/dt + cfke vol + their sum, then
 * line-implicit (Thomas) sweeps along i, j and k, two per direction, with
 * the other directions' increments of the last sweep on the right-hand
 * side; finally v(6:7) += dw.  u holds the residual in planes 6-7 (of ldv).
 * cmt: 1 (1SA, 2JL), 2 (2KO, 2WL), 1.17 (2MT), 1.43 (2Sm) (:82-95).  Wall
 * laws (lparoi != 0) are not provided.  Uses a library workspace. */
int codens_impli2_eqt_3d(const codens_block *b, const double *u, int64_t ldv, const double *dt, double *v,
                         const double *mu, const double *mut, const double *cfke, const double *sn_blk,
                         const double *vol, double cmt, void *stream);

/* impli2_eqt_3d.f90:4 for every block of a layout in one launch per pass (no
 * counterpart in the reference, which calls it block by block from met_num,
 * met_num.f90:697-930): the same results as codens_impli2_eqt_3d called block
 * by block, bit for bit.  blocks: nblk records of one layout (the first at
 * entry 0, none overlapping, all within ldv); sn_all: the metric array of the
 * whole layout (block l's slice at 9 npc).  The block table goes to the
 * device on the first call for a given set of blocks, which must not happen
 * while the stream is captured into a graph. */
int codens_impli2_eqt_3d_batch(int32_t nblk, const codens_block *blocks, const double *u, int64_t ldv,
                               const double *dt, double *v, const double *mu, const double *mut, const double *cfke,
                               const double *sn_all, const double *vol, double cmt, void *stream);

/* met_samut.f90:4 -- SA eddy viscosity mut = rho nu~ fv1 (cv1 = tc[7]). */
int codens_met_samut(const codens_block *b, const double *tc, const double *v, int64_t ldv, const double *mu,
                     double *mut, void *stream);

/* met_kocmut.f90:4 -- k-omega eddy viscosity: rho k/omega (kwsst = 0), or
 * with the SST limiter (kwsst = 1: runs codens_teq_gradv into g9 first, as
 * the reference calls teq_gradv there, rho clamp included; betas = tc[10]). */
int codens_met_kocmut(const codens_block *b, const double *tc, int32_t kwsst, const double *sn_blk,
                      const double *vol, double *v, int64_t ldv, const double *mu, double *mut,
                      const double *dist, double *g9, void *stream);

/* met_smkor.f90:4 -- Wilcox k-omega sources with cross diffusion sigmd
 * max(grad k . grad omega, 0) rho/omega and the strain-rate limiter of the
 * omega production (equatt '2KOR').  fd: gradients of k and omega from
 * codens_met_gradtr; g9, tprod block-local; beta, betas, sigmw = tc[11],
 * tc[10], tc[13].  Writes ts6, ts7 and cfke = 0.18 omega. */
int codens_met_smkor(const codens_block *b, const double *tc, double sigmd, const double *v, int64_t ldv,
                     const double *fd, const double *g9, const double *tprod, double *cfke, double *ts6,
                     double *ts7, void *stream);

/* met_smch.f90:4 -- Chien low-Reynolds k-epsilon sources (equatt '2CH');
 * y+ of a cell from the friction velocity utau[mnpar-1] of its wall point
 * (mnpar: int32 per cell, 0 for none, y+ = 1000), floored at utaumin. */
int codens_met_smch(const codens_block *b, double cke1, double cke2, double utaumin, const double *v, int64_t ldv,
                    const double *mu, const double *dist, const int32_t *mnpar, const double *utau,
                    const double *tprod, double *ts6, double *ts7, void *stream);

/* at_cutke.f90:4 limiters: which = 1 met_cut.f90 (both variables to
 * (epsk, epse) unless both exceed them), which = 2 met_cutsa.f90
 * (nu~ < 0 -> epsk, v(7) = 0). */
int codens_met_cut(const codens_block *b, int32_t which, double epsk, double epse, double *v, int64_t ldv,
                   void *stream);

/* sch_expli.f90:4 -- v += dt/vol * u on the real cells. */
int codens_sch_expli(const codens_block *b, const double *u, double *v, int64_t ldv,
                     const double *dt, const double *vol, void *stream);

/* One stage of the 4-stage explicit scheme built on the single-stage
 * residual (the reference has forward Euler only, sch_expli.f90:53-66):
 * v = v0 - alpha * dt/vol * u on the real cells, planes 0..nvar-1 (u, v0
 * and v share ldv; v may not alias v0 or u).  A step is four calls with
 * alpha = 1/4, 1/3, 1/2, 1, each on the residual of the previous stage's v
 * and the step's fixed dt.  The product and the difference round
 * separately (no fused multiply-add). */
int codens_rk_stage(const codens_block *b, double alpha, const double *v0, const double *u, double *v,
                    int64_t ldv, int32_t nvar, const double *dt, const double *vol, void *stream);

/* implimf_3d.f90:4 -- matrix-free point-Jacobi implicit phase: coefficients,
 * lmx sub-iterations, then v += d.  d (5 arrays of ldv) receives the final
 * increment, zero on the ghosts.  ff is read only when ityprk != 0; dtpas
 * only when kfmg = 3.  Uses a library-owned device workspace.
 * Contract on ps: the pressure of the base flux F(v) is re-derived from v
 * with the zpres formula (zpres.f90:72-73), so ps is not read.  The
 * reference reads ps(n) there (implimf_3d.f90:444-452); the results agree
 * when ps = zpres(v) on the real cells, which holds in atsch_num's stage
 * (zpres runs at its start and v is unchanged until the implicit phase), to
 * rounding (about 3e-14 relative in d).  A caller whose ps differs from
 * zpres(v) gets the increment of zpres(v), not of its ps. */
int codens_implimf_3d(const codens_block *b, const codens_phys *ph, const double *u,
                      const double *dt, double *v, int64_t ldv, double *d, const double *ff,
                      const double *mu, const double *mut, int32_t lmx, const double *sn_blk,
                      const double *vol, double dtpas, int32_t kfmg, int32_t ityprk,
                      const double *ps, const double *cson, void *stream);

/* codens_implimf_3d with no allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as codens_implimf_3d; its scratch (9 doubles per
 * storage entry of the block) is a static device array of the library,
 * sized for blocks of up to 2,400,000 storage entries (172.8 MB per device,
 * held until the process ends); a larger block is refused (CODENS_EINVAL).
 * Every call on a device shares that array: calls that overlap on one device
 * race.  The sub-iteration tile is the forced one (codens_sweep_tile), the
 * one already timed for the block's shape, or the lane model's; no
 * timing launch is made. */
int codens_implimf_3d_noalloc(const codens_block *b, const codens_phys *ph, const double *u, const double *dt,
                              double *v, int64_t ldv, double *d, const double *ff, const double *mu, const double *mut,
                              int32_t lmx, const double *sn_blk, const double *vol, double dtpas, int32_t kfmg,
                              int32_t ityprk, const double *ps, const double *cson, void *stream);

/* FP32 defect-correction form of codens_implimf_3d, with the same arguments
 * and results: the coefficients are computed in FP64 and rounded, and the
 * lmx Jacobi sub-iterations run in FP32 on FP32 copies of v, u, ff and sn;
 * d comes back widened to FP64 and v += d in FP64.  It has no counterpart in
 * the reference (implimf_3d.f90:4 is FP64 throughout).  It is for drivers
 * that iterate to a residual drop: the FP64 residual u fixes the converged
 * state, and the FP32 phase only changes the path to it. */
int codens_implimf_3d_f32(const codens_block *b, const codens_phys *ph, const double *u, const double *dt, double *v,
                          int64_t ldv, double *d, const double *ff, const double *mu, const double *mut, int32_t lmx,
                          const double *sn_blk, const double *vol, double dtpas, int32_t kfmg, int32_t ityprk,
                          const double *ps, const double *cson, void *stream);

/* implimf_eu.f90:6 -- point-Jacobi implicit phase of the Euler equations:
 * coefe = 0.5|V.S| + |S|c, diag = vol (fact + 1/dt) + sum of coefe with
 * fact = 1.5/dt1min when kfmg = 3, flux increments from the momenta and the
 * caller's ps (the z energy increment without its w factor, :312), lmx
 * sub-iterations, then v += d on all nvar planes of v and d (the reference's
 * v(n,:) = v(n,:) + d(n,:), :329; planes 6.. of d are left as passed).  No
 * viscous term, so no mu/mut.  Uses the library's implicit workspace. */
int codens_implimf_eu(const codens_block *b, const codens_phys *ph, const double *u, const double *dt,
                      double *v, int64_t ldv, int32_t nvar, double *d, const double *ff, int32_t lmx,
                      const double *sn_blk, const double *vol, double dt1min, int32_t kfmg, int32_t ityprk,
                      const double *ps, const double *cson, void *stream);

/* Boundary ghosts from face lists (int64 device arrays of flat indices, in
 * the listing order of initis.f90:71-150): ghost cell, adjacent real cell,
 * and for coincident faces the neighbour's adjacent real cell. */

/* rbvc.f90:84-117 -- coincident faces: ghost = average of both real cells. */
int codens_bc_coincident(int64_t nf, const int64_t *ghost, const int64_t *inner,
                         const int64_t *remote, double *v, int64_t ldv, double *ps, double *temp,
                         double *cson, void *stream);
/* clpara.f90:44-58 -- adiabatic no-slip wall. */
int codens_bc_wall(int64_t nf, const int64_t *ghost, const int64_t *inner, double *v, int64_t ldv,
                   double *ps, double *temp, double *cson, void *stream);
/* rbve.f90 + clnrd.f90:60-134 -- non-reflecting far field against the
 * conservative state w0[5] (host array); nrm: 3 device arrays of nf unit
 * normals pointing into the block (initns.f90:56-163). */
int codens_bc_farfield(int64_t nf, const int64_t *ghost, const int64_t *inner, const double *nrm,
                       const double *w0, const codens_phys *ph, double *v, int64_t ldv, double *ps,
                       double *temp, double *cson, void *stream);
/* utnrd.f90:59-105 -- the far-field state w0[5] (host array) that
 * codens_bc_farfield takes: Mach number, incidence angles alpha, beta
 * (degrees; velocity direction (cos a cos b, -sin b, sin a cos b)) and the
 * stagnation state roa1, aa1, pa1 (c_nzst.f90:56-59), as the 'nrd' branch of
 * rbord.f90:300-323 computes it.  Non-rotating frame only (omg = 0, so the
 * node coordinates utnrd reads drop out).  Host computation, no device. */
int codens_utnrd(double mach, double alpha, double beta, double roa1, double aa1, double pa1,
                 const codens_phys *ph, double *w0);
/* residu.f90:131-187 per block and rscpsv.f90:168-192 over all blocks --
 * steadiness norms of an update: for the first neq (1..7) variables,
 * r = (v - u) / dt on the real cells (v the new state, u the old one; the
 * reference passes u = the state before the stage).  out: device array of
 * (nb + 1) * neq * 7 doubles, record (l, e) at ((l * neq) + e) * 7: L1 =
 * (1/N) sum |r|, L2 = sqrt((1/N) sum r^2), Linf = max |r|, then the block l
 * and the (i, j, k) of the last real cell in storage order where |r| =
 * Linf, as residu locates it; record (nb, e) is the aggregate of all blocks
 * (sum N_l L1_l / N, sqrt(sum N_l L2_l^2 / N), max Linf_l, and where that
 * maximum lies).  The reductions run in a fixed order: the result does not
 * depend on timing.  No host synchronisation (the caller reads out when it
 * needs it), so the call can be captured in a HIP graph once its workspace
 * exists.  Uses the library's norms workspace. */
int codens_residu(const codens_block *blocks, int32_t nb, int32_t neq, const double *v, const double *u,
                  int64_t ldv, const double *dt, double *out, void *stream);

/* Non-physical states: adds to *nbad (device int) the real cells whose
 * density or pressure is not positive, or that hold a non-finite value in
 * the first neq (>= 5) variables -- the states on which the reference
 * aborts (zvismo.f90:122-127) or produces NaN.  No counterpart routine in
 * the reference, which checks T only inside zvismo. */
int codens_check_state(const codens_block *blocks, int32_t nb, int32_t neq, const codens_phys *ph,
                       const double *v, int64_t ldv, int32_t *nbad, void *stream);

/* rfve.f90:24-41 -- ghost = adjacent real cell for v (5 planes), p, T and c.
 * atsch_num.f90:450-458 runs it over the non-coincident frontiers in the
 * first cycle of a computation, before rbord; it is what leaves a symmetry
 * ghost's p, T and c at the adjacent cell's first-cycle values. */
int codens_rfve(int64_t nf, const int64_t *ghost, const int64_t *inner, double *v, int64_t ldv, double *ps,
                double *temp, double *cson, void *stream);
/* clsym.f90:56-79 -- symmetry plane (rbord.f90:340): ghost density and
 * energy = adjacent cell's, ghost velocity = adjacent cell's without its
 * normal component; nrm as for codens_bc_farfield.  Like clsym, it writes
 * only v: the ghost's p, T and c keep the values they hold. */
int codens_bc_sym(int64_t nf, const int64_t *ghost, const int64_t *inner, const double *nrm, double *v,
                  int64_t ldv, void *stream);
/* rbte.f90:46-59 -- tq ghost = adjacent real cell. */
int codens_bc_tq_copy(int64_t nf, const int64_t *ghost, const int64_t *inner, double *tq,
                      int64_t ldv, void *stream);
/* rbtc.f90:104-117 -- tq ghost = average of both real cells. */
int codens_bc_tq_coincident(int64_t nf, const int64_t *ghost, const int64_t *inner,
                            const int64_t *remote, double *tq, int64_t ldv, void *stream);

/* met_rbse.f90:4 -- turbulent diffusive flux densities (6 planes of ldv)
 * of boundary ghosts = adjacent real cell. */
int codens_met_rbse(int64_t nf, const int64_t *ghost, const int64_t *inner, double *fd, int64_t ldv,
                    void *stream);
/* met_rbsc.f90:4 -- the same at coincident faces: average of both cells. */
int codens_met_rbsc(int64_t nf, const int64_t *ghost, const int64_t *inner, const int64_t *remote, double *fd,
                    int64_t ldv, void *stream);

/* dua_resro.f90:63-69 / residu.f90:131-187 -- sum over the real cells of all
 * nb blocks of ((a - b)/dt)^2 for one variable (a, b: arrays of ldv).
 * Synchronous; the sum is written to *out. */
int codens_sumsq_rate(const codens_block *blocks, int32_t nb, const double *a, const double *bv,
                      const double *dt, double *out, void *stream);

/* Launch timing (the reference has only a wall clock around `compute flow`,
 * solve.f90:487-515): when enabled, every launch of the main kernels is
 * bracketed by HIP events on its own stream.  Names: "zpres", "chronos",
 * "zvisqc", "roe_limiter_chain", "roe_face_flux", "roe_gather",
 * "implimf_coef", "implimf_sweep", "implimf_update", "sch_expli", "rk_stage".
 * codens_timing_read synchronises on the recorded events and returns the
 * summed duration and launch count of one kernel; reset != 0 clears it. */
int codens_timing_enable(int32_t on);
int codens_timing_read(const char *name, double *ms_total, int64_t *count, int32_t reset);

/* Host-array mirrors for callers whose fields live in host memory (the
 * Fortran drop-in routines in codens_amd/fortran, which keep the dummy
 * arguments of the reference's in-place host routines, e.g. sch_roe.f90:4,
 * implimf_3d.f90:4).  One device buffer per (device, host address) of the
 * array's full `bytes`, grown on demand and kept until the process ends; it
 * has the host array's layout, so device pointers into it use the same
 * offsets.  Only a window moves: `rows` runs of `width` bytes, `pitch` bytes
 * apart, from byte `offset` (one block's slice of each variable).
 * codens_host_mirror returns the buffer in *dev; to_device = 1 copies the
 * window host -> device on `stream`, 2 copies the whole array but only when
 * the buffer is first made (arrays the caller never changes afterwards:
 * metrics and volumes), 0 copies nothing; codens_host_fetch copies the window
 * back and synchronises.  codens_host_bundle packs `count` host arrays of
 * `bytes` each into consecutive slots of one buffer (the 9-array tq and
 * 6-array fd bundles above) with the window [offset, offset+width) of each;
 * codens_host_bundle_fetch copies every slot's window back and synchronises. */
int codens_host_mirror(const void *host, int64_t bytes, int64_t offset, int64_t width, int64_t pitch,
                       int32_t rows, int32_t to_device, void **dev, void *stream);
int codens_host_fetch(void *host, int64_t offset, int64_t width, int64_t pitch, int32_t rows,
                      void *stream);
int codens_host_bundle(const void *const *hosts, int32_t count, int64_t bytes, int64_t offset,
                       int64_t width, int32_t to_device, void **dev, void *stream);
int codens_host_bundle_fetch(const void *const *hosts, int32_t count, int64_t bytes, int64_t offset,
                             int64_t width, void *stream);

/* Residency of the host-array mirrors (no reference counterpart: the
 * drop-in boundary's host/device protocol).  codens_host_residency(0), the
 * default: every window moves in before a call's kernels and back after
 * them, as above.  codens_host_residency(1): the library records, per host
 * array, the windows whose device copy is current and those whose device
 * copy is newer than the host's; a window moving in whose device copy is
 * current is not copied, and a window moving back is not copied but becomes
 * device-newer.  The host then calls codens_host_sync(a) before its own code
 * reads array a (device-newer windows of a come back) and
 * codens_host_modified(a) after its own code writes a (a's device copies are
 * forgotten; CODENS_EINVAL while a has device-newer windows).
 * codens_host_sync_all brings back every device-newer window; turning
 * residency off does the same and forgets every window.  A stage whose host
 * code reads and writes none of its arrays moves nothing over PCIe.
 * codens_host_traffic: bytes moved host -> device and device -> host by the
 * mirrors since the library was loaded.  The arrays are keyed by the address
 * of their first element. */
int codens_host_residency(int32_t on, int32_t *previous, void *stream);
int codens_host_sync(const void *host, void *stream);
int codens_host_sync_all(void *stream);
int codens_host_modified(const void *host);
int codens_host_traffic(int64_t *h2d, int64_t *d2h);

/* ---- Menter k-omega model (met_num.f90:449-456, 542-546, 815-850) ----
 * mc: the model constants betae, okappa, sigme1, sigma1, beta1, wsig1,
 * sigme2, sigma2, beta2, wsig2 (module modeleturb, atlecdon.f90:255-278).
 * Walls: nfaces walls, wall f owning lines[face_start[f] .. face_start[f+1])
 * (face_start in host memory, lines and mnpar in device memory); the walls
 * run one after another in list order, the lines of one wall in parallel
 * (they are distinct cell columns).  Cells a line would reach outside
 * [0, nfrac) are skipped. */

/* met_bord.f90:4 (with met_rbvc.f90:30-65 and met_rbve.f90:26-36) -- the
 * ghost values of the two transported turbulence variables (v planes 5 and
 * 6, 0-based) on one frontier list (ghost g, adjacent real cell in, and for
 * coincident frontiers the real cell across the interface rm), by kind:
 *   0 coincident ('rc'): 0.5 (in + rm), the met_rbvc average;
 *   1 wall ('para') with kparoi = 0: 0 and 0;
 *   2 wall ('para') with kparoi = 1: 0 and the adjacent cell's v7;
 *   3 far field ('nrd'): rokinf and roeinf;
 *   4 symmetry ('sym'): the adjacent cell's values.
 * met_num.f90:145-151 calls it first in the turbulence stage: coincident
 * frontiers, then every frontier in listing order.  in may be NULL for
 * kinds 1 and 3, rm for kinds other than 0. */
int codens_met_bord(int32_t kind, int64_t nf, const int64_t *g, const int64_t *in, const int64_t *rm,
                    double rokinf, double roeinf, double *v, int64_t ldv, void *stream);

/* met_parko.f90:4 -- omega wall condition of the k-omega models (kparoi = 1,
 * met_num.f90:153-159): per wall point (lines[face_start[w]..face_start[w+1])
 * of wall w, as for codens_met_mtcorf1; ghost[p] the point's ghost cell),
 * rho omega = 6 mu / (beta1 d^2) in the wall cell; in the next min(10,
 * m3tot) cells of its line whose wall point (mnpar, 1-based) lies in the
 * wall's window, the value blended with the current one over cells 2..4
 * (pond = (d - d_2) / (d_4 - d_2)) and floored by 6 mu / (beta1 d^2); the
 * ghost takes the wall cell's value.  beta1 = 0.075 for '2KO'.  face_start
 * is host memory; walls run in order.  The reference stops when a line has
 * fewer than 10 cells (met_parko.f90:45-51); the caller checks that. */
int codens_met_parko(int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines, const int64_t *ghost,
                     const int32_t *mnpar, const double *mu, const double *dist, double beta1, double *v, int64_t ldv,
                     void *stream);

/* met_mtcorf1.f90:4 -- wall correction of the blending function frac. */
int codens_met_mtcorf1(int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines, int32_t imxclko,
                       const int32_t *mnpar, double *frac, int64_t nfrac, void *stream);

/* met_smmt.f90:4 -- Menter sources: gkgo (block-local) and frac (kfracom =
 * 0: 0 and 1; else grad k . grad omega from fd and tanh(zeta^4), then
 * codens_met_mtcorf1 over the walls), then ts6, ts7 and cfke = 0.18 omega
 * with the frac-blended constants.  fd: gradients of k and omega (6 planes
 * of ldv, read for kfracom != 0 only); tprod block-local. */
int codens_met_smmt(const codens_block *b, const double *mc, int32_t kfracom, int32_t imxclko, const double *v,
                    int64_t ldv, const double *mu, const double *mut, const double *dist, const double *fd,
                    const double *tprod, double *cfke, double *gkgo, double *frac, double *ts6, double *ts7,
                    int32_t nfaces, const int64_t *face_start, const codens_wall_line *lines, const int32_t *mnpar,
                    void *stream);

/* met_smmtr.f90:4 -- as codens_met_smmt with kfracom = 1 (the blending
 * function always, gkgo kept internal) and Durbin's realizability limit:
 * ts7's production divided by min(1, (1/sqrt(0.09)) / (S/omega)), S from
 * the velocity gradients g9 (block-local, 9 x nnn). */
int codens_met_smmtr(const codens_block *b, const double *mc, int32_t imxclko, const double *v, int64_t ldv,
                     const double *mu, const double *mut, const double *dist, const double *fd, const double *g9,
                     const double *tprod, double *cfke, double *frac, double *ts6, double *ts7, int32_t nfaces,
                     const int64_t *face_start, const codens_wall_line *lines, const int32_t *mnpar, void *stream);

/* met_fludmt.f90:4 -- fd(1:3) *= mu + mut sigme, fd(4:6) *= mu + mut sigma
 * on indc(i1,j1,k1)..indc(i2-1,j2-1,k2-1), the sigmas blended with frac when
 * kfracom != 0 (frac read only then). */
int codens_met_fludmt(const codens_block *b, const double *mc, int32_t kfracom, const double *mu, const double *mut,
                      const double *frac, double *fd, int64_t ldv, void *stream);

/* met_komut.f90:4 -- Menter eddy viscosity: rho k / omega (komsst = 0), or
 * with the SST limiter (komsst = 1: codens_teq_gradv into g9 first, as the
 * reference calls teq_gradv there). */
int codens_met_komut(const codens_block *b, double betae, int32_t komsst, const double *sn_blk, const double *vol,
                     double *v, int64_t ldv, const double *mu, double *mut, const double *dist, double *g9,
                     void *stream);

/* met_smklsas.f90:4 -- codens_met_smkl with Menter's scale-adaptive term:
 * ts7 -= qsi2 P l^3 / (k L_vK^2) where 1 - tanh((8 rtilde)^3) > 0.8, L_vK
 * from the Laplacians of u and v (met_laplaciens, scheme ischema) of the
 * gradients g9 as given. */
int codens_met_smklsas(const codens_block *b, const double *kc, int32_t ischema, const double *sn_blk,
                       const double *vol, const double *v, int64_t ldv, const double *mu, const double *mut,
                       const double *dist, const double *fd, const double *g9, const double *tprod, double *cfke,
                       double *ts6, double *ts7, void *stream);

/* met_smke.f90:4 -- k-epsilon sources: met_exgr on the three rows of the
 * velocity gradients g9 (block-local, 9 x nnn: their ghost layers are
 * written, as the reference does), met_bark and met_bare into bark and bare
 * (block-local), then the variant selected by equatt(4:4): 0 met_smkes
 * (' ' or 'S'), 1 met_smker ('R'), 2 met_smkec ('C', reads cson), 3
 * met_smkesas ('L', Laplacians by met_laplaciens with scheme ischema).
 * kc = cke1, cke2, cmu; lparoi = 1 zeroes bark, bare and the f2 damping.
 * tprod block-local; cfke, ts6 (qcxts5), ts7 (qcyts6) storage arrays. */
int codens_met_smke(const codens_block *b, const double *kc, int32_t variant, int32_t lparoi, int32_t ischema,
                    const double *sn_blk, const double *vol, const double *v, int64_t ldv, const double *mu,
                    const double *mut, const double *tprod, double *g9, const double *cson, double *bark,
                    double *bare, double *cfke, double *ts6, double *ts7, void *stream);

/* met_kemut.f90:4 -- k-epsilon eddy viscosity: form 0 Jones-Launder
 * ('2JL', kesst = 0), 1 with the SST limiter (kesst = 1: codens_teq_gradv
 * into g9 first, as the reference calls teq_gradv there; reads dist), 2
 * Launder-Sharma ('2LS'); fmu = 1 when lparoi is 1 or 2. */
int codens_met_kemut(const codens_block *b, double cmu, int32_t form, int32_t lparoi, const double *sn_blk,
                     const double *vol, double *v, int64_t ldv, const double *mu, double *mut, const double *dist,
                     double *g9, void *stream);

/* sch_roe.f90:4 for every block of a layout in one launch per pass (no
 * counterpart in the reference, which loops over blocks, atsch_num.f90:621):
 * the same results as codens_sch_roe called block by block.  blocks: nblk
 * records of one layout (the first at entry 0, all within ldv); sn_all: the
 * metric array of the whole layout (block l's slice at 9 npc).  The table of
 * blocks goes to the device on the first call for a given set of blocks,
 * which must not happen while the stream is captured into a graph. */
int codens_sch_roe_batch(int32_t nblk, const codens_block *blocks, const codens_phys *ph, const codens_scheme *sc,
                         int32_t ityprk, double *u, const double *v, int64_t ldv, double *ff, const double *tq,
                         const double *sn_all, const double *ps, void *stream);

/* implimf_3d.f90:4 for every block of a layout in one launch per pass (no
 * counterpart in the reference, which calls implimf_3d block by block,
 * atsch_num.f90:1040): the same results as codens_implimf_3d called block by
 * block, bit for bit.  blocks: nblk records of one layout (the first at
 * entry 0, none overlapping, all within ldv); sn_all: the metric array of the
 * whole layout (block l's slice at 9 npc).  The block table goes to the
 * device on the first call for a given set of blocks, which must not happen
 * while the stream is captured into a graph. */
int codens_implimf_3d_batch(int32_t nblk, const codens_block *blocks, const codens_phys *ph, const double *u,
                            const double *dt, double *v, int64_t ldv, double *d, const double *ff, const double *mu,
                            const double *mut, int32_t lmx, const double *sn_all, const double *vol, double dtpas,
                            int32_t kfmg, int32_t ityprk, const double *cson, void *stream);

/* Selects one of 16 independent sets of the library's device workspaces for
 * the entry points called after it (no counterpart in the reference, which
 * runs blocks one after another).  For callers that run blocks
 * concurrently on several streams: each stream its own set, so that no two
 * concurrent calls share scratch memory.  Set 0 is the default. */
int codens_workspace_set(int32_t set);

/* Tile shape of the tiled Jacobi sub-iteration of codens_implimf_3d,
 * codens_implimf_eu and codens_implimf_3d_f32 (no counterpart in the
 * reference): -1 (default) times the 64 x 4, 32 x 8 and 16 x 16 shapes on
 * the first call for each block shape and keeps the fastest, 0 forces
 * 64 x 4, 1 forces 32 x 8, 2 forces 16 x 16 (the all-blocks entry points
 * hold the first two: there a forced 16 x 16 takes the lane model's choice
 * between them: launched lanes times each shape's measured time per lane).
 * All give identical results.  *previous (may be NULL) receives the mode in
 * force before the call. */
int codens_sweep_tile(int32_t mode, int32_t *previous);

/* Kernel arrangement of codens_sch_roe and codens_sch_roe_euler (no
 * counterpart in the reference): 1 (default) one fused kernel that stages
 * an i-j tile of the state in LDS, marches k, evaluates every face once and
 * accumulates the residual of its cells; 0 one face-flux pass per direction
 * into a face array, then a gather; 2 as 1, with codens_sch_roe_batch's
 * fused pass launched persistent: about two workgroups per compute unit
 * take the tiles of every block from a queue (a device counter per
 * workspace set, zeroed when made and left at zero by each launch, so the
 * launch can be captured into a graph after a first eager call).  All
 * evaluate the same face values and sum them in the same order.
 * *previous (may be NULL) receives the arrangement in force before the
 * call. */
int codens_residual_path(int32_t fused, int32_t *previous);

/* The tile shape used for a block (eu: implimf_eu, fp32: the FP32 form):
 * 0 or 1, or -1 if not chosen yet. */
int codens_sweep_tile_of(const codens_block *b, int32_t eu, int32_t fp32, int32_t *tile);

/* ---- Differentiation of the block residual (no counterpart in the
 * reference, which has no adjoint).  The residual is that of
 * codens_sch_roe (sch_roe.f90:4) for laminar Navier-Stokes with unlimited
 * MUSCL (scheme ilim = 0; other values are rejected) after codens_zpres,
 * codens_zvismo, the wall ghosts (codens_bc_wall), the far-field ghosts
 * (codens_bc_farfield, normals and state winf held fixed), codens_zvisqc and
 * the ghost stresses (codens_bc_tq_copy), as a function of the real cells of
 * v (5 x ldv), the block metrics sn (9 x nnn) and vol (ldv).  Only the real
 * cells of u are defined. */

/* Tangent-linear: u and its directional derivative udot for the direction
 * (vdot, sndot, voldot). */
int codens_adj_residual_jvp(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                            const double *v, const double *vdot, int64_t ldv, const double *sn, const double *sndot,
                            const double *vol, const double *voldot, const double *winf, int64_t nwall,
                            const int64_t *wall_ghost, const int64_t *wall_inner, int64_t nfar,
                            const int64_t *far_ghost, const int64_t *far_inner, const double *far_nrm, double *u,
                            double *udot, void *stream);

/* Adjoint: for the adjoint ubar of u (real cells), vbar = (du/dv)^T ubar on
 * the real cells (5 x ldv, the block's entries overwritten, ghosts zero),
 * snbar = (du/dsn)^T ubar (9 x nnn) and volbar = (du/dvol)^T ubar (ldv, the
 * block's entries). */
int codens_adj_residual_vjp(const codens_block *b, const codens_phys *ph, const codens_scheme *sc,
                            const double *v, int64_t ldv, const double *sn, const double *vol, const double *winf,
                            int64_t nwall, const int64_t *wall_ghost, const int64_t *wall_inner, int64_t nfar,
                            const int64_t *far_ghost, const int64_t *far_inner, const double *far_nrm,
                            const double *ubar, double *vbar, double *snbar, double *volbar, void *stream);

/* Mesh metrics of one block from its node coordinates x, y, z (storage
 * arrays, nodes at the cells' flat indices, one ghost node layer on every
 * side): extmhg.f90:88-125 (ex1 = 1, ex2 = 0) first overwrites the ghost
 * nodes with copies of the nearest real nodes (k layers, then i, then j);
 * then norm.f90:102-286 gives sn_blk (9 x nnn, the face area vectors of
 * every node's low i, j, k faces, zero where a face leaves the storage) and
 * svol.f90:90-149 the volumes vol (indexed from npc; the last storage layer
 * of each direction gets 0).  Operations in the reference's order without
 * contraction. */
int codens_mesh_metrics(const codens_block *b, double *x, double *y, double *z, double *sn_blk, double *vol,
                        void *stream);

/* Adjoint of the mesh metrics (norm.f90 face vectors sg_d (d1 x d2) of the
 * half-diagonals, svol.f90 volumes) with respect to the node coordinates
 * x, y, z (storage arrays, nodes at the cells' flat indices): xbar, ybar,
 * zbar (block-local, nnn) from snbar (9 x nnn) and volbar (the block's
 * entries of an ldv array, indexed from npc).  sg: the orientation factor
 * +-0.5 of each direction.  Ghost nodes are treated as independent nodes;
 * folding them onto the real nodes they copy is left to the caller. */
int codens_adj_metric_vjp(const codens_block *b, const double *x, const double *y, const double *z,
                          const double *sg, const double *snbar, const double *volbar, double *xbar, double *ybar,
                          double *zbar, void *stream);

/* ---- On-disk formats of the reference.  Host memory and file I/O only, no
 * GPU work: a device-resident driver copies its arrays out before writing
 * and in after reading.  Files are Fortran unformatted sequential (each
 * record: its byte length as a 4-byte integer, the data, the length again,
 * little endian), one record per variable and block, i fastest, then j, k.
 * A record is limited to 2 GiB (CODENS_EINVAL beyond). */

/* writda.f90:4 (restart, written by svfw.f90:58 after chdacc 'cccc'): for
 * each block in order, records over its real cells ii1..ii2-1, jj1..jj2-1,
 * kk1..kk2-1: rho, rho u, rho v, rho w, rho E (v planes 0..4), then mut when
 * ns != 0 (equat(1:2) = 'ns'), then v planes 5 and 6 when nturb = 2
 * (equat(6:7) = 'ke', or equat(2:4) = 'res' with ip60 = 7).  The file is
 * created or truncated (the reference rewinds for its first block). */
int codens_restart_write(const char *path, const codens_block *blocks, int32_t nb, int32_t ns, int32_t nturb,
                         const double *v, int64_t ldv, const double *mut);

/* readda.f90:4 (read by infw.f90:69): the records of codens_restart_write
 * into the real cells of v (and mut); every other entry is left as it was.
 * CODENS_EIO when the file ends early or a record's length does not match
 * its block. */
int codens_restart_read(const char *path, const codens_block *blocks, int32_t nb, int32_t ns, int32_t nturb,
                        double *v, int64_t ldv, double *mut);

/* writdg.f90:4: for each block, records x, y, z over its nodes ii1..ii2,
 * jj1..jj2, kk1..kk2; node (i, j, k) of block l at the block's storage entry
 * of (i, j, k) (npn = npc), arrays of ld entries. */
int codens_mesh_write(const char *path, const codens_block *blocks, int32_t nb, int64_t ld, const double *x,
                      const double *y, const double *z);

/* readdg.f90:4: the records of codens_mesh_write into the nodes of x, y, z;
 * other entries untouched. */
int codens_mesh_read(const char *path, const codens_block *blocks, int32_t nb, int64_t ld, double *x, double *y,
                     double *z);

#ifdef __cplusplus
}
#endif
#endif /* CODENS_H */
