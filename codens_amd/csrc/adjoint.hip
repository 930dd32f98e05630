// Discrete tangent-linear (JVP) and adjoint (VJP) of the block residual of
// sch_roe on gfx950, for laminar Navier-Stokes with unlimited kappa-MUSCL
// (ilim = 0), adiabatic walls (clpara) and characteristic far fields (clnrd):
//
//   u(v; sn, vol) = sch_roe(v, ps, tq(v), sn)   after
//   zpres and zvismo on the real cells, the wall and far-field ghosts, and
//   zvisqc (Green-Gauss gradients, stresses, ghost stresses = adjacent cell),
//
// with the far-field normals and the free-stream state held fixed.  The
// minmod form (ilim = 1) is not covered: its chained slope ratios
// (sch_roe.f90:146-194) make each face's limiter a continued fraction over
// the whole grid line, and the reference build itself amplifies 1e-15
// changes of the state to 5e-3 within 10 stages there.
//
// Every piece of the residual is written once, templated on the scalar:
// double gives the values, Dual (value, tangent) gives exact directional
// derivatives.  The JVP runs the chain on Dual planes.  The VJP walks the
// chain backwards; where a piece is nonlinear (the face fluxes, the ghost
// rules, the thermodynamics, the stresses) the transposed product of its
// local Jacobian with the incoming adjoint is formed from one Dual pass per
// input of that piece, so it is exact to rounding, and the adjoints are
// scattered to the cells with FP64 atomic adds.
//
// Arithmetic restates oracle_flux.c / oracle_visc.c (sch_roe.f90:322-578,
// zpres.f90:66-79, zvismo.f90:90-104, clpara.f90:44-58, clnrd.f90:60-134,
// zgrad2.f90:148-485, zfluto.f90:104-121) in the same operation order, so
// the double instantiation reproduces the residual to rounding.
#include "common.hpp"

#pragma clang fp contract(off)

using namespace codens;

namespace {

// ---------------------------------------------------------------------------
// dual numbers
struct Dual {
  double v, d;
};
__device__ inline Dual mk(double v, double d = 0.) { return Dual{v, d}; }
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ inline Dual operator/(double a, Dual b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ inline Dual sqrt(Dual a) {
  double s = ::sqrt(a.v);
  return {s, 0.5 * a.d / s};
}
__device__ inline Dual fabs(Dual a) { return a.v < 0. ? -a : a; }
__device__ inline double sqrt(double a) { return ::sqrt(a); }
__device__ inline double fabs(double a) { return ::fabs(a); }
__device__ inline double val(double a) { return a; }
__device__ inline double val(Dual a) { return a.v; }
__device__ inline double tan_(Dual a) { return a.d; }
template <typename T>
__device__ inline T zro();
template <>
__device__ inline double zro<double>() { return 0.; }
template <>
__device__ inline Dual zro<Dual>() { return Dual{0., 0.}; }

// a plane pair (values, tangents); T = double reads the values only
template <typename T>
struct Pl;
template <>
struct Pl<double> {
  const double *v;
  __device__ double operator[](long i) const { return v[i]; }
};
template <>
struct Pl<Dual> {
  const double *v, *d;
  __device__ Dual operator[](long i) const { return Dual{v[i], d[i]}; }
};
template <typename T>
struct Wp;
template <>
struct Wp<double> {
  double *v;
  __device__ void set(long i, double x) const { v[i] = x; }
};
template <>
struct Wp<Dual> {
  double *v, *d;
  __device__ void set(long i, Dual x) const {
    v[i] = x.v;
    d[i] = x.d;
  }
};

struct APhys {
  double gam, gam1, pinfl, pr, cp, usrey, asuth;
  int iflu;
};

// zpres.f90:66-79 then zvismo.f90:90-104 (mut = 0)
template <typename T>
__device__ inline void thermo(const APhys &ph, const T w[5], T &p, T &tt, T &mu) {
  T rhoe = w[4] - 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) / w[0];
  p = ph.gam1 * (rhoe - ph.pinfl);
  tt = ph.gam * p / w[0];
  if (ph.iflu == 1)
    mu = ph.usrey * tt * sqrt(tt) * (1. + ph.asuth) / (tt + ph.asuth);
  else
    mu = tt * 0. + ph.usrey;
}

// Ghost state of a boundary face as a function of the adjacent real cell:
// conservative state wg (read by the boundary flux), primitive state qg
// (rho, u, v, w, p: read by the MUSCL reconstruction) and temperature tg
// (read by the gradients).
template <typename T>
__device__ inline void ghost_prims(const T wg[5], T pg, T qg[5]) {
  qg[0] = wg[0];
  qg[1] = wg[1] / wg[0];
  qg[2] = wg[2] / wg[0];
  qg[3] = wg[3] / wg[0];
  qg[4] = pg;
}

// clpara.f90:44-58
template <typename T>
__device__ inline void wall_ghost(const APhys &ph, const T vi[5], T wg[5], T qg[5], T &tg) {
  T rhoe1 = vi[4] - 0.5 * (vi[1] * vi[1] + vi[2] * vi[2] + vi[3] * vi[3]) / vi[0];
  T p, mu;
  thermo(ph, vi, p, tg, mu);
  wg[0] = vi[0];
  wg[1] = vi[1] * 0.;
  wg[2] = vi[1] * 0.;
  wg[3] = vi[1] * 0.;
  wg[4] = rhoe1;
  ghost_prims(wg, p, qg);
}

// rbve.f90 copy then clnrd.f90:60-134 against the state w0 (inward unit
// normal n)
template <typename T>
__device__ inline void far_ghost(const APhys &ph, const double *w0, double nx, double ny, double nz, const T vi[5],
                                 T wg[5], T qg[5], T &tg) {
  const double gam = ph.gam, gam1 = ph.gam1, pinfl = ph.pinfl;
  T ro0 = vi[0];
  T ps0 = gam1 * (vi[4] - 0.5 * (vi[1] * vi[1] + vi[2] * vi[2] + vi[3] * vi[3]) / vi[0] - pinfl);
  T roc0 = sqrt(gam * ps0 * ro0);
  T roqn0 = vi[1] * nx + vi[2] * ny + vi[3] * nz;
  double epsm = .5 + copysign(.5, val(roc0 - roqn0));
  double eps0 = .5 + copysign(.5, -val(roqn0));
  double epsp = .5 + copysign(.5, val(-roc0 - roqn0));
  double rod = w0[0];
  double qxd = w0[1] / rod, qyd = w0[2] / rod, qzd = w0[3] / rod;
  double pd = gam1 * (w0[4] - .5 * rod * (qxd * qxd + qyd * qyd + qzd * qzd) - pinfl);
  double qnd = qxd * nx + qyd * ny + qzd * nz;
  double qtxd = qxd - qnd * nx, qtyd = qyd - qnd * ny, qtzd = qzd - qnd * nz;
  T ros = vi[0];
  T qxs = vi[1] / ros, qys = vi[2] / ros, qzs = vi[3] / ros;
  T p_s = gam1 * (vi[4] - .5 * ros * (qxs * qxs + qys * qys + qzs * qzs) - pinfl);
  T qns = qxs * nx + qys * ny + qzs * nz;
  T qtxs = qxs - qns * nx, qtys = qys - qns * ny, qtzs = qzs - qns * nz;
  T qtx = eps0 * qtxs + (1. - eps0) * qtxd;
  T qty = eps0 * qtys + (1. - eps0) * qtyd;
  T qtz = eps0 * qtzs + (1. - eps0) * qtzd;
  T am = epsm * (p_s - roc0 * qns) + (1. - epsm) * (pd - roc0 * qnd);
  T ap = epsp * (p_s + roc0 * qns) + (1. - epsp) * (pd + roc0 * qnd);
  T qn = 0.5 * (ap - am) / roc0;
  T p = 0.5 * (ap + am);
  T bs = (p - p_s) * (ro0 * ro0) / (roc0 * roc0) + ros;
  T b0 = (p - pd) * (ro0 * ro0) / (roc0 * roc0) + rod;
  T ro = eps0 * bs + (1. - eps0) * b0;
  wg[0] = ro;
  wg[1] = ro * (qtx + qn * nx);
  wg[2] = ro * (qty + qn * ny);
  wg[3] = ro * (qtz + qn * nz);
  wg[4] = p / gam1 + pinfl + 0.5 * (wg[1] * wg[1] + wg[2] * wg[2] + wg[3] * wg[3]) / ro;
  tg = gam * p / ro;
  ghost_prims(wg, p, qg);
}

// zfluto.f90:104-121 with mut = 0; g: the 12 scaled gradients
template <typename T>
__device__ inline void stresses(const APhys &ph, const T g[12], T mu, T tq[9]) {
  const double ds3 = 2. / 3.;
  T xx = g[0], xy = g[1], xz = g[2], yx = g[3], yy = g[4], yz = g[5], zx = g[6], zy = g[7], zz = g[8];
  T mt = mu + 0.;
  tq[0] = ds3 * mt * (2 * xx - yy - zz);
  tq[3] = ds3 * mt * (2 * yy - xx - zz);
  tq[5] = ds3 * mt * (2 * zz - xx - yy);
  tq[1] = mt * (xy + yx);
  tq[2] = mt * (xz + zx);
  tq[4] = mt * (yz + zy);
  T kq = ph.cp * (mu / ph.pr + 0.);
  tq[6] = kq * g[9];
  tq[7] = kq * g[10];
  tq[8] = kq * g[11];
}

// zgrad2.f90:463-485: scale 0.5 / vol, guarded for vol <= 0
template <typename T>
__device__ inline T grad_scale(T vol) {
  if (!signbit(val(vol))) return .5 / (vol + DBL_MIN);  // vols = vol
  return vol * 0. + .5 / (0.00001 + DBL_MIN);            // vols = eps
}

// Roe flux of one face (2x the numerical flux . S), sch_roe.f90:327-473
template <typename T>
__device__ void roe_face_t(const APhys &ph, const T s[3], const T L[5], const T R[5], const T tl[9], const T tr[9],
                           T f[5]) {
  const double gam = ph.gam, gam1 = ph.gam1, pinfl = ph.pinfl;
  T rhol = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  T rhor = R[0], ur = R[1], vr = R[2], wr = R[3], prr = R[4];
  const double igam1 = 1. / gam1;
  T s0 = s[0], s1 = s[1], s2 = s[2];
  T cnds = sqrt(s0 * s0 + s1 * s1 + s2 * s2), icn = 1. / cnds;
  T nx = s0 * icn, ny = s1 * icn, nz = s2 * icn;
  T irl = 1. / rhol, irr = 1. / rhor;
  T al2 = gam * pl * irl, ar2 = gam * prr * irr;
  T q2l = ul * ul + vl * vl + wl * wl, q2r = ur * ur + vr * vr + wr * wr;
  T hl = al2 * igam1 + 0.5 * q2l, hr = ar2 * igam1 + 0.5 * q2r;
  T el = (pl * igam1 + pinfl) * irl + 0.5 * q2l;
  T er = (prr * igam1 + pinfl) * irr + 0.5 * q2r;
  T gd = sqrt(rhor * irl), gd1 = 1. / (1. + gd), gd2 = gd * gd1;
  T rhom = sqrt(rhol * rhor), rhomi = 1. / rhom;
  T um = gd1 * ul + gd2 * ur, vm = gd1 * vl + gd2 * vr, wm = gd1 * wl + gd2 * wr;
  T hm = gd1 * hl + gd2 * hr;
  T vitm2 = 0.5 * (um * um + vm * vm + wm * wm);
  T am = sqrt(gam1 * (hm - vitm2)), ami = 1. / am;
  T am2i = ami * ami;
  T vn = um * nx + vm * ny + wm * nz;
  T rhoiam = rhom * ami, rhoami = ami * rhomi;
  T l1 = fabs(vn * cnds), l4 = fabs(vn * cnds + am * cnds), l5 = fabs(vn * cnds - am * cnds);
  T dw0 = rhor - rhol;
  T dw1 = rhor * ur - rhol * ul;
  T dw2 = rhor * vr - rhol * vl;
  T dw3 = rhor * wr - rhol * wl;
  T dw4 = rhor * er - rhol * el;
  T g = gam1 * vitm2 * am2i;
  T qdw = gam1 * (um * dw1 + vm * dw2 + wm * dw3 - dw4);
  T a0 = (1. - g) * nx * dw0 - (vm * nz - wm * ny) * rhomi * dw0 + nx * am2i * qdw + (nz * dw2 - ny * dw3) * rhomi;
  T a1 = (1. - g) * ny * dw0 - (wm * nx - um * nz) * rhomi * dw0 + ny * am2i * qdw + (nx * dw3 - nz * dw1) * rhomi;
  T a2 = (1. - g) * nz * dw0 - (um * ny - vm * nx) * rhomi * dw0 + nz * am2i * qdw + (ny * dw1 - nx * dw2) * rhomi;
  T mdw = (nx * dw1 + ny * dw2 + nz * dw3) * rhomi;
  T cdw = gam1 * rhoami * (vitm2 * dw0 - um * dw1 - vm * dw2 - wm * dw3 + dw4);
  T a3 = cdw - vn * rhomi * dw0 + mdw;
  T a4 = cdw + vn * rhomi * dw0 - mdw;
  a0 = a0 * l1; a1 = a1 * l1; a2 = a2 * l1; a3 = a3 * l4; a4 = a4 * l5;
  T h = 0.5 * rhoiam;
  T di0 = nx * a0 + ny * a1 + nz * a2 + h * (a3 + a4);
  T di1 = um * nx * a0 + (um * ny - rhom * nz) * a1 + (um * nz + rhom * ny) * a2 +
          h * ((um + nx * am) * a3 + (um - nx * am) * a4);
  T di2 = (vm * nx + rhom * nz) * a0 + vm * ny * a1 + (vm * nz - rhom * nx) * a2 +
          h * ((vm + ny * am) * a3 + (vm - ny * am) * a4);
  T di3 = (wm * nx - rhom * ny) * a0 + (wm * ny + rhom * nx) * a1 + wm * nz * a2 +
          h * ((wm + nz * am) * a3 + (wm - nz * am) * a4);
  T di4 = (vitm2 * nx + rhom * (vm * nz - wm * ny)) * a0 + (vitm2 * ny + rhom * (wm * nx - um * nz)) * a1 +
          (vitm2 * nz + rhom * (um * ny - vm * nx)) * a2 + h * ((hm + am * vn) * a3 + (hm - am * vn) * a4);
  T dfxx = rhor * (ur * ur) + prr - pinfl - tr[0] + rhol * (ul * ul) + pl - pinfl - tl[0];
  T dfxy = rhor * ur * vr - tr[1] + rhol * ul * vl - tl[1];
  T dfxz = rhor * ur * wr - tr[2] + rhol * ul * wl - tl[2];
  T dfyy = rhor * (vr * vr) + prr - pinfl - tr[3] + rhol * (vl * vl) + pl - pinfl - tl[3];
  T dfyz = rhor * vr * wr - tr[4] + rhol * vl * wl - tl[4];
  T dfzz = rhor * (wr * wr) + prr - pinfl - tr[5] + rhol * (wl * wl) + pl - pinfl - tl[5];
  T Er = rhor * er + prr - pinfl, El = rhol * el + pl - pinfl;
  T dfex = Er * ur + El * ul - (tr[0] * ur + tr[1] * vr + tr[2] * wr) - (tl[0] * ul + tl[1] * vl + tl[2] * wl) -
           tr[6] - tl[6];
  T dfey = Er * vr + El * vl - (tr[1] * ur + tr[3] * vr + tr[4] * wr) - (tl[1] * ul + tl[3] * vl + tl[4] * wl) -
           tr[7] - tl[7];
  T dfez = Er * wr + El * wl - (tr[2] * ur + tr[4] * vr + tr[5] * wr) - (tl[2] * ul + tl[4] * vl + tl[5] * wl) -
           tr[8] - tl[8];
  f[0] = (rhor * ur + rhol * ul) * s0 + (rhor * vr + rhol * vl) * s1 + (rhor * wr + rhol * wl) * s2 - di0;
  f[1] = dfxx * s0 + dfxy * s1 + dfxz * s2 - di1;
  f[2] = dfxy * s0 + dfyy * s1 + dfyz * s2 - di2;
  f[3] = dfxz * s0 + dfyz * s1 + dfzz * s2 - di3;
  f[4] = dfex * s0 + dfey * s1 + dfez * s2 - di4;
}

// Physical flux of a ghost state on a boundary face, Navier-Stokes form
// (sch_roe.f90:497-578; oracle_flux.c phys_flux with rho_div = w[0])
template <typename T>
__device__ void phys_flux_t(const APhys &ph, const T w[5], T p, const T t[9], const T s[3], T f[5]) {
  const double pinfl = ph.pinfl;
  T fxx = w[1] * (w[1] / w[0]) + p - pinfl - t[0];
  T fxy = w[2] * (w[1] / w[0]) - t[1];
  T fxz = w[3] * (w[1] / w[0]) - t[2];
  T fyy = w[2] * (w[2] / w[0]) + p - pinfl - t[3];
  T fyz = w[3] * (w[2] / w[0]) - t[4];
  T fzz = w[3] * (w[3] / w[0]) + p - pinfl - t[5];
  T fex = ((w[4] + p - pinfl - t[0]) * w[1] - t[1] * w[2] - t[2] * w[3]) / w[0] - t[6];
  T fey = ((w[4] + p - pinfl - t[3]) * w[2] - t[1] * w[1] - t[4] * w[3]) / w[0] - t[7];
  T fez = ((w[4] + p - pinfl - t[5]) * w[3] - t[2] * w[1] - t[4] * w[2]) / w[0] - t[8];
  f[0] = w[1] * s[0] + w[2] * s[1] + w[3] * s[2];
  f[1] = fxx * s[0] + fxy * s[1] + fxz * s[2];
  f[2] = fxy * s[0] + fyy * s[1] + fyz * s[2];
  f[3] = fxz * s[0] + fyz * s[1] + fzz * s[2];
  f[4] = fex * s[0] + fey * s[1] + fez * s[2];
}

// ---------------------------------------------------------------------------
// Workspace planes (block-local index m, nnn entries each; JVP tangents in a
// second set).  Q: primitive state of every real and ghost cell; TT: T;
// MU: mu of the real cells; W: conservative ghost state; G: raw Green-Gauss
// sums (12); C0: gradient scale; TQ: stresses and heat flux (9); H: face
// fluxes (3 directions x 5).
constexpr int P_Q = 0, P_TT = 5, P_MU = 6, P_W = 7, P_G = 12, P_C0 = 24, P_TQ = 25, P_H = 34, NPLANE = 49;

struct AdjArgs {
  Blk b;
  APhys ph;
  double muscl, xk;
  long ldv;
  const double *v, *vd;      // state (5 x ldv) and its tangent
  const double *sn, *snd;    // block metrics (9 x nnn) and tangent
  const double *vol, *vold;  // ldv
  const double *w0;          // far-field state (5)
  double *ws, *wd;           // planes and tangent planes
};

template <typename T>
__device__ inline Pl<T> plane(const AdjArgs &a, int p);
template <>
__device__ inline Pl<double> plane<double>(const AdjArgs &a, int p) {
  return {a.ws + (long)p * a.b.nnn};
}
template <>
__device__ inline Pl<Dual> plane<Dual>(const AdjArgs &a, int p) {
  return {a.ws + (long)p * a.b.nnn, a.wd + (long)p * a.b.nnn};
}
template <typename T>
__device__ inline Wp<T> wplane(const AdjArgs &a, int p);
template <>
__device__ inline Wp<double> wplane<double>(const AdjArgs &a, int p) {
  return {a.ws + (long)p * a.b.nnn};
}
template <>
__device__ inline Wp<Dual> wplane<Dual>(const AdjArgs &a, int p) {
  return {a.ws + (long)p * a.b.nnn, a.wd + (long)p * a.b.nnn};
}
template <typename T>
__device__ inline T ldx(const double *v, const double *d, long i);
template <>
__device__ inline double ldx<double>(const double *v, const double *, long i) {
  return v[i];
}
template <>
__device__ inline Dual ldx<Dual>(const double *v, const double *d, long i) {
  return Dual{v[i], d[i]};
}

template <typename T>
__device__ inline void load_v(const AdjArgs &a, long n, T w[5]) {
#pragma unroll
  for (int e = 0; e < 5; e++) w[e] = ldx<T>(a.v, a.vd, e * a.ldv + n);
}
template <typename T>
__device__ inline T load_sn(const AdjArgs &a, long m, int d, int c) {
  long o = ((long)c * 3 + d) * a.b.nnn + m;
  return ldx<T>(a.sn, a.snd, o);
}

// real cell of a 1-D launch over the real cells, i fastest
__device__ inline bool real_cell(const Blk &b, long t, int c[3]) {
  long nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  if (t >= nx * ny * nz) return false;
  c[0] = b.i1 + (int)(t % nx);
  c[1] = b.j1 + (int)((t / nx) % ny);
  c[2] = b.k1 + (int)(t / (nx * ny));
  return true;
}
// face of direction d (along index a1..a2, real transverse range), i fastest
__device__ inline bool face_of(const Blk &b, int d, long t, int c[3]) {
  long n0 = b.i2 - b.i1 + (d == 0), n1 = b.j2 - b.j1 + (d == 1), n2 = b.k2 - b.k1 + (d == 2);
  if (t >= n0 * n1 * n2) return false;
  c[0] = b.i1 + (int)(t % n0);
  c[1] = b.j1 + (int)((t / n0) % n1);
  c[2] = b.k1 + (int)(t / (n0 * n1));
  return true;
}
__device__ inline int along(const int c[3], int d) { return c[d]; }

// ---------------------------------------------------------------------------
// forward pieces (T = double: values; T = Dual: values and tangents)

template <typename T>
__global__ void __launch_bounds__(256) k_adj_thermo(AdjArgs a) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 256L + threadIdx.x, c)) return;
  long n = a.b.idx(c[0], c[1], c[2]), m = n - a.b.npc;
  T w[5], p, tt, mu;
  load_v(a, n, w);
  thermo(a.ph, w, p, tt, mu);
  T q[5] = {w[0], w[1] / w[0], w[2] / w[0], w[3] / w[0], p};
#pragma unroll
  for (int e = 0; e < 5; e++) wplane<T>(a, P_Q + e).set(m, q[e]);
  wplane<T>(a, P_TT).set(m, tt);
  wplane<T>(a, P_MU).set(m, mu);
}

template <typename T, int FAR>
__global__ void __launch_bounds__(256) k_adj_ghost(AdjArgs a, long nf, const int64_t *__restrict__ g,
                                                   const int64_t *__restrict__ in, const double *__restrict__ nrm) {
  long f = blockIdx.x * 256L + threadIdx.x;
  if (f >= nf) return;
  long mg = g[f] - a.b.npc;
  T vi[5], wg[5], qg[5], tg;
  load_v(a, in[f], vi);
  if (FAR)
    far_ghost(a.ph, a.w0, nrm[f], nrm[nf + f], nrm[2 * nf + f], vi, wg, qg, tg);
  else
    wall_ghost(a.ph, vi, wg, qg, tg);
#pragma unroll
  for (int e = 0; e < 5; e++) {
    wplane<T>(a, P_W + e).set(mg, wg[e]);
    wplane<T>(a, P_Q + e).set(mg, qg[e]);
  }
  wplane<T>(a, P_TT).set(mg, tg);
}

// value of gradient quantity q (u, v, w, T) at block-local m
template <typename T>
__device__ inline T gq(const AdjArgs &a, int q, long m) {
  return q < 3 ? plane<T>(a, P_Q + 1 + q)[m] : plane<T>(a, P_TT)[m];
}

// Raw Green-Gauss sums of one real cell in zgrad2's accumulation order
// (directions k, j, i; per direction the interior low face, the interior
// high face, then the boundary faces) and the scale.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad(AdjArgs a) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 256L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long m = b.idx(c[0], c[1], c[2]) - b.npc;
  T G[12];
#pragma unroll
  for (int x = 0; x < 12; x++) G[x] = zro<T>();
  for (int d = 2; d >= 0; d--) {
    long inc = b.inc(d);
    int al = c[d], a1 = b.lo(d), a2 = b.hi(d);
    bool lo_in = al > a1, hi_in = al < a2 - 1;
    T slo[3], shi[3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      slo[cc] = load_sn<T>(a, m, d, cc);
      shi[cc] = load_sn<T>(a, m + inc, d, cc);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      T here = gq<T>(a, q, m);
      if (lo_in) {
        T fv = here + gq<T>(a, q, m - inc);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) G[3 * q + cc] = G[3 * q + cc] - fv * slo[cc];
      }
      if (hi_in) {
        T fv = gq<T>(a, q, m + inc) + here;
#pragma unroll
        for (int cc = 0; cc < 3; cc++) G[3 * q + cc] = G[3 * q + cc] + fv * shi[cc];
      }
    }
    if (!lo_in) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T gv = gq<T>(a, q, m - inc);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) G[3 * q + cc] = G[3 * q + cc] - slo[cc] * 2. * gv;
      }
    }
    if (!hi_in) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T gv = gq<T>(a, q, m + inc);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) G[3 * q + cc] = G[3 * q + cc] + shi[cc] * 2. * gv;
      }
    }
  }
#pragma unroll
  for (int x = 0; x < 12; x++) wplane<T>(a, P_G + x).set(m, G[x]);
  long n = m + b.npc;
  wplane<T>(a, P_C0).set(m, grad_scale(ldx<T>(a.vol, a.vold, n)));
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_tq(AdjArgs a) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 256L + threadIdx.x, c)) return;
  long m = a.b.idx(c[0], c[1], c[2]) - a.b.npc;
  T g[12], tq[9];
  T c0 = plane<T>(a, P_C0)[m];
#pragma unroll
  for (int x = 0; x < 12; x++) g[x] = plane<T>(a, P_G + x)[m] * c0;
  stresses(a.ph, g, plane<T>(a, P_MU)[m], tq);
#pragma unroll
  for (int x = 0; x < 9; x++) wplane<T>(a, P_TQ + x).set(m, tq[x]);
}

// rbte.f90: ghost stresses = adjacent real cell
template <typename T>
__global__ void __launch_bounds__(256) k_adj_tq_ghost(AdjArgs a, long nf, const int64_t *__restrict__ g,
                                                      const int64_t *__restrict__ in) {
  long f = blockIdx.x * 256L + threadIdx.x;
  if (f >= nf) return;
  long mg = g[f] - a.b.npc, mi = in[f] - a.b.npc;
#pragma unroll
  for (int x = 0; x < 9; x++) wplane<T>(a, P_TQ + x).set(mg, plane<T>(a, P_TQ + x)[mi]);
}

// MUSCL coefficients of q(a-2), q(a-1), q(a), q(a+1) in the left and right
// states of the face between a-1 and a (sch_roe_euler.f90:143-171)
template <typename T>
__device__ inline void muscl_lr(double muscl, double xk, const T q0, const T q1, const T q2, const T q3, T &L,
                                T &R) {
  L = q1 + 0.25 * muscl * ((1. - xk) * (q1 - q0) + (1. + xk) * (q2 - q1));
  R = q2 - 0.25 * muscl * ((1. + xk) * (q2 - q1) + (1. - xk) * (q3 - q2));
}

// face flux H (0.5 x Roe flux on interior faces, the ghost's physical flux
// on boundary faces) of direction d at face index m
template <typename T>
__device__ inline void face_h(const AdjArgs &a, int d, const int c[3], T H[5]) {
  const Blk &b = a.b;
  long inc = b.inc(d), m = b.idx(c[0], c[1], c[2]) - b.npc;
  int al = c[d], a1 = b.lo(d), a2 = b.hi(d);
  T s[3];
#pragma unroll
  for (int cc = 0; cc < 3; cc++) s[cc] = load_sn<T>(a, m, d, cc);
  if (al > a1 && al < a2) {
    T L[5], R[5], tl[9], tr[9], f[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      Pl<T> q = plane<T>(a, P_Q + e);
      muscl_lr(a.muscl, a.xk, q[m - 2 * inc], q[m - inc], q[m], q[m + inc], L[e], R[e]);
    }
#pragma unroll
    for (int x = 0; x < 9; x++) {
      tl[x] = plane<T>(a, P_TQ + x)[m - inc];
      tr[x] = plane<T>(a, P_TQ + x)[m];
    }
    roe_face_t(a.ph, s, L, R, tl, tr, f);
#pragma unroll
    for (int e = 0; e < 5; e++) H[e] = f[e] * 0.5;
  } else {
    long mg = al == a1 ? m - inc : m;
    T w[5], t[9];
#pragma unroll
    for (int e = 0; e < 5; e++) w[e] = plane<T>(a, P_W + e)[mg];
#pragma unroll
    for (int x = 0; x < 9; x++) t[x] = plane<T>(a, P_TQ + x)[mg];
    phys_flux_t(a.ph, w, plane<T>(a, P_Q + 4)[mg], t, s, H);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_faces(AdjArgs a, int d) {
  int c[3];
  if (!face_of(a.b, d, blockIdx.x * 256L + threadIdx.x, c)) return;
  long m = a.b.idx(c[0], c[1], c[2]) - a.b.npc;
  T H[5];
  face_h(a, d, c, H);
#pragma unroll
  for (int e = 0; e < 5; e++) wplane<T>(a, P_H + 5 * d + e).set(m, H[e]);
}

// u = sum over directions of H(high face) - H(low face) on the real cells
template <typename T>
__global__ void __launch_bounds__(256) k_adj_gather(AdjArgs a, double *u, double *ud) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 256L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
#pragma unroll
  for (int e = 0; e < 5; e++) {
    Pl<T> h0 = plane<T>(a, P_H + e), h1 = plane<T>(a, P_H + 5 + e), h2 = plane<T>(a, P_H + 10 + e);
    T r = h0[m + 1] - h0[m];
    r = r + (h1[m + b.nid] - h1[m]);
    r = r + (h2[m + b.nijd] - h2[m]);
    u[e * a.ldv + n] = val(r);
    if constexpr (sizeof(T) == sizeof(Dual)) ud[e * a.ldv + n] = tan_(r);
  }
}

// ---------------------------------------------------------------------------
// adjoint pieces.  Adjoint planes (block-local, nnn each): QB primitive
// state (5), WB conservative ghost state (5), VB gradient quantities u, v,
// w, T (4), TQB stresses (9), GB raw gradient sums (12), MUB mu.
constexpr int B_Q = 0, B_W = 5, B_V = 10, B_TQ = 14, B_G = 23, B_MU = 35, NBPLANE = 36;

struct AdjBar {
  double *bw;                 // adjoint planes
  const double *ub;           // adjoint of u (5 x ldv)
  double *vb, *snb, *volb;    // results: v (5 x ldv), sn (9 x nnn), vol (ldv)
};

__device__ inline double *bpl(const AdjArgs &a, const AdjBar &B, int p) { return B.bw + (long)p * a.b.nnn; }

// Transposed Jacobian of the face flux against Hb, one Dual pass per input.
__global__ void __launch_bounds__(128) k_adjb_faces(AdjArgs a, AdjBar B, int d) {
  int c[3];
  if (!face_of(a.b, d, blockIdx.x * 128L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long inc = b.inc(d), m = b.idx(c[0], c[1], c[2]) - b.npc, n = m + b.npc;
  int al = c[d], a1 = b.lo(d), a2 = b.hi(d);
  const double *ub = B.ub;
  const long ldv = a.ldv;
  double s[3];
#pragma unroll
  for (int cc = 0; cc < 3; cc++) s[cc] = load_sn<double>(a, m, d, cc);
  double sb[3] = {0., 0., 0.};
  if (al > a1 && al < a2) {
    // u(m) has -H(m), u(m-inc) has +H(m); H = 0.5 f
    double fb[5];
#pragma unroll
    for (int e = 0; e < 5; e++) fb[e] = 0.5 * (ub[e * ldv + n - inc] - ub[e * ldv + n]);
    double q[4][5], L[5], R[5], tl[9], tr[9];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      const double *qp = a.ws + (long)(P_Q + e) * b.nnn;
#pragma unroll
      for (int k = 0; k < 4; k++) q[k][e] = qp[m + (k - 2) * inc];
      muscl_lr(a.muscl, a.xk, q[0][e], q[1][e], q[2][e], q[3][e], L[e], R[e]);
    }
#pragma unroll
    for (int x = 0; x < 9; x++) {
      tl[x] = a.ws[(long)(P_TQ + x) * b.nnn + m - inc];
      tr[x] = a.ws[(long)(P_TQ + x) * b.nnn + m];
    }
    double xb[31];
#pragma unroll 1
    for (int k = 0; k < 31; k++) {
      Dual sD[3], LD[5], RD[5], tlD[9], trD[9], fD[5];
#pragma unroll
      for (int x = 0; x < 3; x++) sD[x] = mk(s[x], k == 28 + x);
#pragma unroll
      for (int x = 0; x < 5; x++) {
        LD[x] = mk(L[x], k == x);
        RD[x] = mk(R[x], k == 5 + x);
      }
#pragma unroll
      for (int x = 0; x < 9; x++) {
        tlD[x] = mk(tl[x], k == 10 + x);
        trD[x] = mk(tr[x], k == 19 + x);
      }
      roe_face_t(a.ph, sD, LD, RD, tlD, trD, fD);
      double acc = 0.;
#pragma unroll
      for (int e = 0; e < 5; e++) acc += fb[e] * fD[e].d;
      xb[k] = acc;
    }
    // MUSCL: L = cl0 q0 + cl1 q1 + cl2 q2, R = cr1 q1 + cr2 q2 + cr3 q3
    const double c4 = 0.25 * a.muscl;
    const double cl0 = -c4 * (1. - a.xk), cl1 = 1. + c4 * ((1. - a.xk) - (1. + a.xk)), cl2 = c4 * (1. + a.xk);
    const double cr1 = c4 * (1. + a.xk), cr2 = 1. - c4 * ((1. + a.xk) - (1. - a.xk)), cr3 = -c4 * (1. - a.xk);
#pragma unroll
    for (int e = 0; e < 5; e++) {
      double *qb = bpl(a, B, B_Q + e);
      double Lb = xb[e], Rb = xb[5 + e];
      atomicAdd(qb + m - 2 * inc, cl0 * Lb);
      atomicAdd(qb + m - inc, cl1 * Lb + cr1 * Rb);
      atomicAdd(qb + m, cl2 * Lb + cr2 * Rb);
      atomicAdd(qb + m + inc, cr3 * Rb);
    }
#pragma unroll
    for (int x = 0; x < 9; x++) {
      atomicAdd(bpl(a, B, B_TQ + x) + m - inc, xb[10 + x]);
      atomicAdd(bpl(a, B, B_TQ + x) + m, xb[19 + x]);
    }
#pragma unroll
    for (int x = 0; x < 3; x++) sb[x] = xb[28 + x];
  } else {
    bool lo = al == a1;
    long mg = lo ? m - inc : m;
    double hb[5];
#pragma unroll
    for (int e = 0; e < 5; e++) hb[e] = lo ? -ub[e * ldv + n] : ub[e * ldv + n - inc];
    double w[5], t[9], p = a.ws[(long)(P_Q + 4) * b.nnn + mg];
#pragma unroll
    for (int e = 0; e < 5; e++) w[e] = a.ws[(long)(P_W + e) * b.nnn + mg];
#pragma unroll
    for (int x = 0; x < 9; x++) t[x] = a.ws[(long)(P_TQ + x) * b.nnn + mg];
    double xb[18];
#pragma unroll 1
    for (int k = 0; k < 18; k++) {
      Dual wD[5], tD[9], sD[3], fD[5];
#pragma unroll
      for (int x = 0; x < 5; x++) wD[x] = mk(w[x], k == x);
      Dual pD = mk(p, k == 5);
#pragma unroll
      for (int x = 0; x < 9; x++) tD[x] = mk(t[x], k == 6 + x);
#pragma unroll
      for (int x = 0; x < 3; x++) sD[x] = mk(s[x], k == 15 + x);
      phys_flux_t(a.ph, wD, pD, tD, sD, fD);
      double acc = 0.;
#pragma unroll
      for (int e = 0; e < 5; e++) acc += hb[e] * fD[e].d;
      xb[k] = acc;
    }
    // one writer per ghost for W and p; stresses of the ghost through TQB
#pragma unroll
    for (int e = 0; e < 5; e++) atomicAdd(bpl(a, B, B_W + e) + mg, xb[e]);
    atomicAdd(bpl(a, B, B_Q + 4) + mg, xb[5]);
#pragma unroll
    for (int x = 0; x < 9; x++) atomicAdd(bpl(a, B, B_TQ + x) + mg, xb[6 + x]);
#pragma unroll
    for (int x = 0; x < 3; x++) sb[x] = xb[15 + x];
  }
#pragma unroll
  for (int cc = 0; cc < 3; cc++) atomicAdd(B.snb + ((long)cc * 3 + d) * b.nnn + m, sb[cc]);
}

// ghost stresses were copies of the adjacent cell
__global__ void __launch_bounds__(256) k_adjb_tq_ghost(AdjArgs a, AdjBar B, long nf, const int64_t *__restrict__ g,
                                                       const int64_t *__restrict__ in) {
  long f = blockIdx.x * 256L + threadIdx.x;
  if (f >= nf) return;
  long mg = g[f] - a.b.npc, mi = in[f] - a.b.npc;
#pragma unroll
  for (int x = 0; x < 9; x++) atomicAdd(bpl(a, B, B_TQ + x) + mi, bpl(a, B, B_TQ + x)[mg]);
}

// stresses of a real cell: adjoints of the raw gradient sums, of mu and of
// vol (through the scale), one Dual pass per input (12 sums, mu, vol)
__global__ void __launch_bounds__(128) k_adjb_tq(AdjArgs a, AdjBar B) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 128L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long m = b.idx(c[0], c[1], c[2]) - b.npc, n = m + b.npc;
  double G[12], tb[9], mu = a.ws[(long)P_MU * b.nnn + m], vol = a.vol[n];
#pragma unroll
  for (int x = 0; x < 12; x++) G[x] = a.ws[(long)(P_G + x) * b.nnn + m];
#pragma unroll
  for (int x = 0; x < 9; x++) tb[x] = bpl(a, B, B_TQ + x)[m];
#pragma unroll 1
  for (int k = 0; k < 14; k++) {
    Dual g[12], tq[9];
    Dual c0 = grad_scale(mk(vol, k == 13));
#pragma unroll
    for (int x = 0; x < 12; x++) g[x] = mk(G[x], k == x) * c0;
    stresses(a.ph, g, mk(mu, k == 12), tq);
    double acc = 0.;
#pragma unroll
    for (int x = 0; x < 9; x++) acc += tb[x] * tq[x].d;
    if (k < 12)
      bpl(a, B, B_G + k)[m] = acc;
    else if (k == 12)
      bpl(a, B, B_MU)[m] = acc;
    else
      B.volb[n] += acc;
  }
}

// Green-Gauss sums of one face, transposed: adjoints of the face's
// quantities (both cells, or twice the ghost) and of its metric
__global__ void __launch_bounds__(256) k_adjb_grad(AdjArgs a, AdjBar B, int d) {
  int c[3];
  if (!face_of(a.b, d, blockIdx.x * 256L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long inc = b.inc(d), m = b.idx(c[0], c[1], c[2]) - b.npc;
  int al = c[d], a1 = b.lo(d), a2 = b.hi(d);
  double s[3], sb[3] = {0., 0., 0.};
#pragma unroll
  for (int cc = 0; cc < 3; cc++) s[cc] = load_sn<double>(a, m, d, cc);
  if (al > a1 && al < a2) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double fv = gq<double>(a, q, m) + gq<double>(a, q, m - inc), fvb = 0.;
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        const double *gb = bpl(a, B, B_G + 3 * q + cc);
        double tb = gb[m - inc] - gb[m];
        fvb += tb * s[cc];
        sb[cc] += tb * fv;
      }
      atomicAdd(bpl(a, B, B_V + q) + m, fvb);
      atomicAdd(bpl(a, B, B_V + q) + m - inc, fvb);
    }
  } else {
    bool lo = al == a1;
    long mg = lo ? m - inc : m, mr = lo ? m : m - inc;
    double sgn = lo ? -1. : 1.;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double gv = gq<double>(a, q, mg), gvb = 0.;
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        double tb = sgn * bpl(a, B, B_G + 3 * q + cc)[mr];
        gvb += tb * s[cc] * 2.;
        sb[cc] += tb * 2. * gv;
      }
      atomicAdd(bpl(a, B, B_V + q) + mg, gvb);
    }
  }
#pragma unroll
  for (int cc = 0; cc < 3; cc++) atomicAdd(B.snb + ((long)cc * 3 + d) * b.nnn + m, sb[cc]);
}

// ghost rules transposed: adjoints of the ghost's W, q and T onto the
// adjacent cell's conservative state, one Dual pass per input
template <int FAR>
__global__ void __launch_bounds__(128) k_adjb_ghost(AdjArgs a, AdjBar B, long nf, const int64_t *__restrict__ g,
                                                    const int64_t *__restrict__ in, const double *__restrict__ nrm) {
  long f = blockIdx.x * 128L + threadIdx.x;
  if (f >= nf) return;
  const Blk &b = a.b;
  long mg = g[f] - b.npc, ni = in[f];
  double wb[5], qb[5], tb;
#pragma unroll
  for (int e = 0; e < 5; e++) {
    wb[e] = bpl(a, B, B_W + e)[mg];
    qb[e] = bpl(a, B, B_Q + e)[mg];
  }
#pragma unroll
  for (int q = 0; q < 3; q++) qb[1 + q] += bpl(a, B, B_V + q)[mg];
  tb = bpl(a, B, B_V + 3)[mg];
  double vi[5];
  load_v(a, ni, vi);
#pragma unroll 1
  for (int k = 0; k < 5; k++) {
    Dual vD[5], wg[5], qg[5], tg;
#pragma unroll
    for (int x = 0; x < 5; x++) vD[x] = mk(vi[x], k == x);
    if (FAR)
      far_ghost(a.ph, a.w0, nrm[f], nrm[nf + f], nrm[2 * nf + f], vD, wg, qg, tg);
    else
      wall_ghost(a.ph, vD, wg, qg, tg);
    double acc = tb * tg.d;
#pragma unroll
    for (int e = 0; e < 5; e++) acc += wb[e] * wg[e].d + qb[e] * qg[e].d;
    atomicAdd(B.vb + k * a.ldv + ni, acc);
  }
}

// real cells: primitive state, T and mu onto the conservative state
__global__ void __launch_bounds__(128) k_adjb_thermo(AdjArgs a, AdjBar B) {
  int c[3];
  if (!real_cell(a.b, blockIdx.x * 128L + threadIdx.x, c)) return;
  const Blk &b = a.b;
  long n = b.idx(c[0], c[1], c[2]), m = n - b.npc;
  double qb[5], tb = bpl(a, B, B_V + 3)[m], mub = bpl(a, B, B_MU)[m];
#pragma unroll
  for (int e = 0; e < 5; e++) qb[e] = bpl(a, B, B_Q + e)[m];
#pragma unroll
  for (int q = 0; q < 3; q++) qb[1 + q] += bpl(a, B, B_V + q)[m];
  double w[5];
  load_v(a, n, w);
#pragma unroll 1
  for (int k = 0; k < 5; k++) {
    Dual wD[5], p, tt, mu;
#pragma unroll
    for (int x = 0; x < 5; x++) wD[x] = mk(w[x], k == x);
    thermo(a.ph, wD, p, tt, mu);
    Dual q[5] = {wD[0], wD[1] / wD[0], wD[2] / wD[0], wD[3] / wD[0], p};
    double acc = tb * tt.d + mub * mu.d;
#pragma unroll
    for (int e = 0; e < 5; e++) acc += qb[e] * q[e].d;
    B.vb[k * a.ldv + n] += acc;
  }
}

__global__ void k_adj_zero(double *p, long cnt) {
  long i = blockIdx.x * 256L + threadIdx.x;
  if (i < cnt) p[i] = 0.;
}

// np planes of leading dimension ld, entries off .. off+cnt-1
__global__ void k_adj_zero_planes(double *p, long ld, int np, long off, long cnt) {
  long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= cnt * np) return;
  p[(i / cnt) * ld + off + i % cnt] = 0.;
}

// ---------------------------------------------------------------------------
// metric adjoint (norm.f90 half-diagonal cross products, svol.f90 volumes)

struct MArgs {
  Blk b;
  const double *x, *y, *z;
  double sg[3];
  double *xb, *yb, *zb;
  long nid, njd, nkd;
};

// storage offset of node (i, j, k) counted from the block's first entry
__device__ inline long nodeo(const MArgs &M, long i, long j, long k) { return i + M.nid * (j + M.njd * k); }

__device__ inline void corner_ofs(int d, int q, int &di, int &dj, int &dk) {
  // svol.f90 face corners: k faces (0,0,0)(1,0,0)(0,1,0)(1,1,0); j faces
  // (0,0,0)(1,0,0)(0,0,1)(1,0,1); i faces (0,0,0)(0,1,0)(0,0,1)(0,1,1)
  const int K[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
  const int J[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}};
  const int I[4][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}};
  const int(*T)[3] = d == 2 ? K : (d == 1 ? J : I);
  di = T[q][0]; dj = T[q][1]; dk = T[q][2];
}

// the two half-diagonals of a face: d1 = P(e1) - P(0), d2 = P(e2) - P(e3)
__device__ inline void diag_ofs(int d, int e[3][3]) {
  if (d == 2) {
    int t[3][3] = {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}};
    for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) e[a][c] = t[a][c];
  } else if (d == 1) {
    int t[3][3] = {{1, 0, 1}, {0, 0, 1}, {1, 0, 0}};
    for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) e[a][c] = t[a][c];
  } else {
    int t[3][3] = {{0, 1, 1}, {0, 0, 1}, {0, 1, 0}};
    for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) e[a][c] = t[a][c];
  }
}

// volume adjoints onto the six face sums of each cell: vol = (1/12) x
// (s_i(i+1) - s_i(i) + s_j(j+1) - s_j(j) + s_k(k+1) - s_k(k))
__global__ void __launch_bounds__(256) k_madj_vol(MArgs M, const double *__restrict__ volb, double *fsb) {
  long t = blockIdx.x * 256L + threadIdx.x;
  long nc = (M.nid - 1) * (M.njd - 1) * (M.nkd - 1);
  if (t >= nc) return;
  long i = t % (M.nid - 1), j = (t / (M.nid - 1)) % (M.njd - 1), k = t / ((M.nid - 1) * (M.njd - 1));
  long o = nodeo(M, i, j, k), nnn = M.b.nnn;
  double vb = volb[M.b.npc + o] * (0.25 / 3.);
  atomicAdd(fsb + o + 1, vb);
  atomicAdd(fsb + o, -vb);
  atomicAdd(fsb + nnn + o + M.nid, vb);
  atomicAdd(fsb + nnn + o, -vb);
  atomicAdd(fsb + 2 * nnn + o + M.nid * M.njd, vb);
  atomicAdd(fsb + 2 * nnn + o, -vb);
}

// one face of direction d: face sum and cross product transposed onto the
// nodes (snb: the metric adjoint, 9 x nnn, block-local)
__global__ void __launch_bounds__(256) k_madj_face(MArgs M, int d, const double *__restrict__ snb,
                                                   const double *__restrict__ fsb) {
  long t = blockIdx.x * 256L + threadIdx.x;
  long n0 = M.nid - (d != 0), n1 = M.njd - (d != 1), n2 = M.nkd - (d != 2);
  if (t >= n0 * n1 * n2) return;
  long i = t % n0, j = (t / n0) % n1, k = t / (n0 * n1);
  long o = nodeo(M, i, j, k), nnn = M.b.nnn;
  const double *P[3] = {M.x + M.b.npc, M.y + M.b.npc, M.z + M.b.npc};
  double *PB[3] = {M.xb, M.yb, M.zb};
  // sn of this face from the nodes
  int e[3][3];
  diag_ofs(d, e);
  long oe1 = nodeo(M, i + e[0][0], j + e[0][1], k + e[0][2]);
  long oe2 = nodeo(M, i + e[1][0], j + e[1][1], k + e[1][2]);
  long oe3 = nodeo(M, i + e[2][0], j + e[2][1], k + e[2][2]);
  double d1[3], d2[3];
  for (int c = 0; c < 3; c++) {
    d1[c] = P[c][oe1] - P[c][o];
    d2[c] = P[c][oe2] - P[c][oe3];
  }
  double sn[3] = {M.sg[d] * (d1[1] * d2[2] - d1[2] * d2[1]), M.sg[d] * (d1[2] * d2[0] - d1[0] * d2[2]),
                  M.sg[d] * (d1[0] * d2[1] - d1[1] * d2[0])};
  // face sum fs = sum_c (P0 + P1 + P2 + P3)_c sn_c
  double fb = fsb[(long)d * nnn + o];
  double snbar[3];
  long oc[4];
  for (int q = 0; q < 4; q++) {
    int di, dj, dk;
    corner_ofs(d, q, di, dj, dk);
    oc[q] = nodeo(M, i + di, j + dj, k + dk);
  }
  for (int c = 0; c < 3; c++) {
    double cs = ((P[c][oc[0]] + P[c][oc[1]]) + P[c][oc[2]]) + P[c][oc[3]];
    snbar[c] = snb[((long)c * 3 + d) * nnn + o] + fb * cs;
    for (int q = 0; q < 4; q++) atomicAdd(PB[c] + oc[q], fb * sn[c]);
  }
  // sn = sg (d1 x d2): d1b = d2 x tb, d2b = tb x d1 with tb = sg snbar
  double tb[3] = {M.sg[d] * snbar[0], M.sg[d] * snbar[1], M.sg[d] * snbar[2]};
  double d1b[3] = {d2[1] * tb[2] - d2[2] * tb[1], d2[2] * tb[0] - d2[0] * tb[2], d2[0] * tb[1] - d2[1] * tb[0]};
  double d2b[3] = {tb[1] * d1[2] - tb[2] * d1[1], tb[2] * d1[0] - tb[0] * d1[2], tb[0] * d1[1] - tb[1] * d1[0]};
  for (int c = 0; c < 3; c++) {
    atomicAdd(PB[c] + oe1, d1b[c]);
    atomicAdd(PB[c] + o, -d1b[c]);
    atomicAdd(PB[c] + oe2, d2b[c]);
    atomicAdd(PB[c] + oe3, -d2b[c]);
  }
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

int adj_setup(const char *what, const codens_block *cb, const codens_phys *cp, const codens_scheme *sc,
              const double *v, int64_t ldv, const double *sn, const double *vol, const double *winf, long nwall,
              const int64_t *wg, const int64_t *wi, long nfar, const int64_t *fg, const int64_t *fi,
              const double *fn, AdjArgs &a) {
  int rc = codens_check_block(cb, what);
  if (rc) return rc;
  if (!cp || !sc || !v || !sn || !vol || (nfar && (!fg || !fi || !fn || !winf)) || (nwall && (!wg || !wi)) ||
      ldv <= 0 || nwall < 0 || nfar < 0) {
    codens_set_error("%s: null or invalid argument", what);
    return CODENS_EINVAL;
  }
  if (sc->ilim != 0) {
    codens_set_error("%s: the adjoint covers unlimited MUSCL (ilim = 0) only", what);
    return CODENS_EINVAL;
  }
  a.b = make_blk(cb);
  a.ph.gam = cp->gam; a.ph.gam1 = cp->gam1; a.ph.pinfl = cp->pinfl; a.ph.pr = cp->pr; a.ph.cp = cp->cp;
  a.ph.usrey = 1. / cp->reynz; a.ph.asuth = 110.4 / cp->tnz; a.ph.iflu = cp->iflu;
  a.muscl = (double)sc->muscl; a.xk = sc->xk; a.ldv = (long)ldv;
  a.v = v; a.sn = sn; a.vol = vol; a.w0 = winf;
  a.vd = nullptr; a.snd = nullptr; a.vold = nullptr; a.wd = nullptr;
  return CODENS_OK;
}

template <typename T>
void forward_chain(const AdjArgs &a, long nwall, const int64_t *wg, const int64_t *wi, long nfar,
                   const int64_t *fg, const int64_t *fi, const double *fn, hipStream_t st) {
  const Blk &b = a.b;
  long ncell = (long)(b.i2 - b.i1) * (b.j2 - b.j1) * (b.k2 - b.k1);
  hipLaunchKernelGGL(k_adj_thermo<T>, dim3(nblk(ncell, 256)), dim3(256), 0, st, a);
  if (nwall) hipLaunchKernelGGL((k_adj_ghost<T, 0>), dim3(nblk(nwall, 256)), dim3(256), 0, st, a, nwall, wg, wi, fn);
  if (nfar) hipLaunchKernelGGL((k_adj_ghost<T, 1>), dim3(nblk(nfar, 256)), dim3(256), 0, st, a, nfar, fg, fi, fn);
  hipLaunchKernelGGL(k_adj_grad<T>, dim3(nblk(ncell, 256)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_adj_tq<T>, dim3(nblk(ncell, 256)), dim3(256), 0, st, a);
  if (nwall) hipLaunchKernelGGL(k_adj_tq_ghost<T>, dim3(nblk(nwall, 256)), dim3(256), 0, st, a, nwall, wg, wi);
  if (nfar) hipLaunchKernelGGL(k_adj_tq_ghost<T>, dim3(nblk(nfar, 256)), dim3(256), 0, st, a, nfar, fg, fi);
}

long nfaces(const Blk &b, int d) {
  return (long)(b.i2 - b.i1 + (d == 0)) * (b.j2 - b.j1 + (d == 1)) * (b.k2 - b.k1 + (d == 2));
}

}  // namespace

extern "C" int codens_adj_residual_jvp(const codens_block *cb, const codens_phys *cp, const codens_scheme *sc,
                                       const double *v, const double *vdot, int64_t ldv, const double *sn,
                                       const double *sndot, const double *vol, const double *voldot,
                                       const double *winf, int64_t nwall, const int64_t *wall_ghost,
                                       const int64_t *wall_inner, int64_t nfar, const int64_t *far_ghost,
                                       const int64_t *far_inner, const double *far_nrm, double *u, double *udot,
                                       void *stream) {
  const char *what = "codens_adj_residual_jvp";
  AdjArgs a;
  int rc = adj_setup(what, cb, cp, sc, v, ldv, sn, vol, winf, (long)nwall, wall_ghost, wall_inner, (long)nfar,
                     far_ghost, far_inner, far_nrm, a);
  if (rc) return rc;
  if (!vdot || !sndot || !voldot || !u || !udot) {
    codens_set_error("%s: null argument", what);
    return CODENS_EINVAL;
  }
  long nnn = a.b.nnn;
  double *ws = (double *)codens_workspace(sizeof(double) * 2 * NPLANE * nnn, 8);
  if (!ws) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  hipStream_t st = (hipStream_t)stream;
  a.ws = ws; a.wd = ws + NPLANE * nnn;
  a.vd = vdot; a.snd = sndot; a.vold = voldot;
  hipLaunchKernelGGL(k_adj_zero, dim3(nblk(2 * NPLANE * nnn, 256)), dim3(256), 0, st, ws, 2 * NPLANE * nnn);
  forward_chain<Dual>(a, nwall, wall_ghost, wall_inner, nfar, far_ghost, far_inner, far_nrm, st);
  for (int d = 0; d < 3; d++)
    hipLaunchKernelGGL(k_adj_faces<Dual>, dim3(nblk(nfaces(a.b, d), 256)), dim3(256), 0, st, a, d);
  long ncell = (long)(a.b.i2 - a.b.i1) * (a.b.j2 - a.b.j1) * (a.b.k2 - a.b.k1);
  hipLaunchKernelGGL(k_adj_gather<Dual>, dim3(nblk(ncell, 256)), dim3(256), 0, st, a, u, udot);
  return codens_check_launch(what);
}

extern "C" int codens_adj_residual_vjp(const codens_block *cb, const codens_phys *cp, const codens_scheme *sc,
                                       const double *v, int64_t ldv, const double *sn, const double *vol,
                                       const double *winf, int64_t nwall, const int64_t *wall_ghost,
                                       const int64_t *wall_inner, int64_t nfar, const int64_t *far_ghost,
                                       const int64_t *far_inner, const double *far_nrm, const double *ubar,
                                       double *vbar, double *snbar, double *volbar, void *stream) {
  const char *what = "codens_adj_residual_vjp";
  AdjArgs a;
  int rc = adj_setup(what, cb, cp, sc, v, ldv, sn, vol, winf, (long)nwall, wall_ghost, wall_inner, (long)nfar,
                     far_ghost, far_inner, far_nrm, a);
  if (rc) return rc;
  if (!ubar || !vbar || !snbar || !volbar) {
    codens_set_error("%s: null argument", what);
    return CODENS_EINVAL;
  }
  long nnn = a.b.nnn;
  double *ws = (double *)codens_workspace(sizeof(double) * (NPLANE + NBPLANE) * nnn, 8);
  if (!ws) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  hipStream_t st = (hipStream_t)stream;
  a.ws = ws;
  AdjBar B{ws + NPLANE * nnn, ubar, vbar, snbar, volbar};
  hipLaunchKernelGGL(k_adj_zero, dim3(nblk((NPLANE + NBPLANE) * nnn, 256)), dim3(256), 0, st, ws,
                     (NPLANE + NBPLANE) * nnn);
  // results are overwritten on the block's entries
  hipLaunchKernelGGL(k_adj_zero_planes, dim3(nblk(5 * nnn, 256)), dim3(256), 0, st, vbar, (long)ldv, 5,
                     (long)a.b.npc, nnn);
  hipLaunchKernelGGL(k_adj_zero, dim3(nblk(9 * nnn, 256)), dim3(256), 0, st, snbar, 9 * nnn);
  hipLaunchKernelGGL(k_adj_zero_planes, dim3(nblk(nnn, 256)), dim3(256), 0, st, volbar, nnn, 1, (long)a.b.npc,
                     nnn);
  forward_chain<double>(a, nwall, wall_ghost, wall_inner, nfar, far_ghost, far_inner, far_nrm, st);
  long ncell = (long)(a.b.i2 - a.b.i1) * (a.b.j2 - a.b.j1) * (a.b.k2 - a.b.k1);
  for (int d = 0; d < 3; d++)
    hipLaunchKernelGGL(k_adjb_faces, dim3(nblk(nfaces(a.b, d), 128)), dim3(128), 0, st, a, B, d);
  if (nwall)
    hipLaunchKernelGGL(k_adjb_tq_ghost, dim3(nblk(nwall, 256)), dim3(256), 0, st, a, B, (long)nwall, wall_ghost,
                       wall_inner);
  if (nfar)
    hipLaunchKernelGGL(k_adjb_tq_ghost, dim3(nblk(nfar, 256)), dim3(256), 0, st, a, B, (long)nfar, far_ghost,
                       far_inner);
  hipLaunchKernelGGL(k_adjb_tq, dim3(nblk(ncell, 128)), dim3(128), 0, st, a, B);
  for (int d = 0; d < 3; d++)
    hipLaunchKernelGGL(k_adjb_grad, dim3(nblk(nfaces(a.b, d), 256)), dim3(256), 0, st, a, B, d);
  if (nwall)
    hipLaunchKernelGGL(k_adjb_ghost<0>, dim3(nblk(nwall, 128)), dim3(128), 0, st, a, B, (long)nwall, wall_ghost,
                       wall_inner, far_nrm);
  if (nfar)
    hipLaunchKernelGGL(k_adjb_ghost<1>, dim3(nblk(nfar, 128)), dim3(128), 0, st, a, B, (long)nfar, far_ghost,
                       far_inner, far_nrm);
  hipLaunchKernelGGL(k_adjb_thermo, dim3(nblk(ncell, 128)), dim3(128), 0, st, a, B);
  return codens_check_launch(what);
}

extern "C" int codens_adj_metric_vjp(const codens_block *cb, const double *x, const double *y, const double *z,
                                     const double *sg3, const double *snbar, const double *volbar, double *xbar,
                                     double *ybar, double *zbar, void *stream) {
  const char *what = "codens_adj_metric_vjp";
  int rc = codens_check_block(cb, what);
  if (rc) return rc;
  if (!x || !y || !z || !sg3 || !snbar || !volbar || !xbar || !ybar || !zbar) {
    codens_set_error("%s: null argument", what);
    return CODENS_EINVAL;
  }
  MArgs M;
  M.b = make_blk(cb);
  M.x = x; M.y = y; M.z = z;
  for (int d = 0; d < 3; d++) M.sg[d] = sg3[d];
  M.xb = xbar; M.yb = ybar; M.zb = zbar;
  M.nid = M.b.nid; M.njd = M.b.nijd / M.b.nid; M.nkd = M.b.nnn / M.b.nijd;
  long nnn = M.b.nnn;
  double *fsb = (double *)codens_workspace(sizeof(double) * 3 * nnn, 9);
  if (!fsb) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_adj_zero, dim3(nblk(3 * nnn, 256)), dim3(256), 0, st, fsb, 3 * nnn);
  for (double *pb : {xbar, ybar, zbar}) hipLaunchKernelGGL(k_adj_zero, dim3(nblk(nnn, 256)), dim3(256), 0, st, pb, nnn);
  long nc = (M.nid - 1) * (M.njd - 1) * (M.nkd - 1);
  hipLaunchKernelGGL(k_madj_vol, dim3(nblk(nc, 256)), dim3(256), 0, st, M, volbar, fsb);
  for (int d = 0; d < 3; d++) {
    long nf = (M.nid - (d != 0)) * (M.njd - (d != 1)) * (M.nkd - (d != 2));
    hipLaunchKernelGGL(k_madj_face, dim3(nblk(nf, 256)), dim3(256), 0, st, M, d, snbar, fsb);
  }
  return codens_check_launch(what);
}
