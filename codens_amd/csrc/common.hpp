// Shared device-side definitions of the gfx950 residual-and-update kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include <float.h>
#include <stdint.h>

#include "../../include/codens.h"

namespace codens {

// Block extents in the form the kernels use: flat entry of (i, j, k) is
// npc + (i-id1) + (j-jd1)*nid + (k-kd1)*nijd (crdms.f90:126-153).
struct Blk {
  long npc, nid, nijd, nnn;
  int id1, jd1, kd1;
  int i1, i2, j1, j2, k1, k2;  // node extents: real cells i1..i2-1

  __host__ __device__ inline long idx(int i, int j, int k) const {
    return npc + (i - id1) + (long)(j - jd1) * nid + (long)(k - kd1) * nijd;
  }
  __host__ __device__ inline long inc(int d) const { return d == 0 ? 1 : (d == 1 ? nid : nijd); }
  __host__ __device__ inline int lo(int d) const { return d == 0 ? i1 : (d == 1 ? j1 : k1); }
  __host__ __device__ inline int hi(int d) const { return d == 0 ? i2 : (d == 1 ? j2 : k2); }
  __host__ __device__ inline bool real(int i, int j, int k) const {
    return i >= i1 && i < i2 && j >= j1 && j < j2 && k >= k1 && k < k2;
  }
};

struct Phys {
  double gam, gam1, pinfl, pr, prt, cp, reynz, tnz;
  double igam1;  // 1 / gam1
  int iflu;
};

inline Blk make_blk(const codens_block *b) {
  Blk k;
  k.npc = b->npc;
  k.nid = b->id2 - b->id1 + 1;
  k.nijd = k.nid * (long)(b->jd2 - b->jd1 + 1);
  k.nnn = b->nnn;
  k.id1 = b->id1; k.jd1 = b->jd1; k.kd1 = b->kd1;
  k.i1 = b->ii1; k.i2 = b->ii2; k.j1 = b->jj1; k.j2 = b->jj2; k.k1 = b->kk1; k.k2 = b->kk2;
  return k;
}

inline Phys make_phys(const codens_phys *p) {
  Phys q;
  q.gam = p->gam; q.gam1 = p->gam1; q.pinfl = p->pinfl; q.pr = p->pr; q.prt = p->prt;
  q.cp = p->cp; q.reynz = p->reynz; q.tnz = p->tnz; q.iflu = p->iflu;
  q.igam1 = 1. / p->gam1;
  return q;
}

// 1/x for a finite, normal, nonzero x by Newton's iteration for the
// reciprocal, r <- r + r (1 - x r), from the hardware estimate (each step
// squares the relative error, so two steps end within an ulp).  Five VALU
// instructions in place of the eleven of a correctly rounded division.  On
// 2^24 random operands of exponents -1000..1000 it equals the correctly
// rounded 1/x on every one; it is not meant for 0, infinities or operands
// whose reciprocal is subnormal.
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.);
  r = fma(r, e, r);
  e = fma(-x, r, 1.);
  return fma(r, e, r);
}

// sn(m, d, c) of a block's metric slice (snorm.f90:65-71)
__device__ inline double snv(const double *__restrict__ sn, long nnn, long m, int d, int c) {
  return sn[((long)c * 3 + d) * nnn + m];
}

// Launch shape for kernels over the real cells: 64 (i) x 4 (j) threads,
// one k plane per grid z.
constexpr int TX = 64, TY = 4;
// Work-group coordinates of a 3-D launch, remapped so that each of the 8 XCDs
// (which receive work-groups round robin in launch order) owns a contiguous
// share of the grid: j and k neighbours of a tile then sit in the same L2.
// With rev, each XCD runs its share in reverse order: a kernel that follows
// one over the same arrays then begins on the data that kernel touched last,
// still in the last-level cache.
__device__ inline dim3 xcd_block(bool rev = false) {
  unsigned gx = gridDim.x, gy = gridDim.y, nwg = gx * gy * gridDim.z;
  unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  unsigned x = hw % 8u, q = nwg / 8u, r = nwg % 8u;
  unsigned lg = x * q + (x < r ? x : r) + hw / 8u;
  if (rev) lg = nwg - 1 - lg;
  return dim3(lg % gx, (lg / gx) % gy, lg / (gx * gy));
}

inline dim3 cell_grid(const Blk &b) {
  int nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  return dim3((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz);
}

// The non-real (ghost) entries of a block's storage, in this order: the k
// planes outside k1..k2-1; the j rows outside j1..j2-1 of the real planes;
// the i cells outside i1..i2-1 of the real rows.  shell_count(b) is their
// number, shell_entry(b, t) the storage index of the t-th.  A kernel that
// clears the ghosts of an array whose real cells another kernel writes
// then touches only these (an eighth of the block's storage for 64^3
// cells, more for small blocks).
struct Shell {
  long nid, nijd, njd, nkd;
  int li0, li1, lj0, lj1, lk0, lk1;  // real cells, storage-local, half-open
  long na, nb, nc;                    // entries of the three parts
};

__host__ __device__ inline Shell make_shell(const Blk &b) {
  Shell s;
  s.nid = b.nid; s.nijd = b.nijd; s.njd = b.nijd / b.nid; s.nkd = b.nnn / b.nijd;
  s.li0 = b.i1 - b.id1; s.li1 = b.i2 - b.id1;
  s.lj0 = b.j1 - b.jd1; s.lj1 = b.j2 - b.jd1;
  s.lk0 = b.k1 - b.kd1; s.lk1 = b.k2 - b.kd1;
  s.na = (s.lk0 + (s.nkd - s.lk1)) * s.nijd;
  s.nb = (long)(s.lk1 - s.lk0) * (s.lj0 + (s.njd - s.lj1)) * s.nid;
  s.nc = (long)(s.lk1 - s.lk0) * (s.lj1 - s.lj0) * (s.li0 + (s.nid - s.li1));
  return s;
}

__host__ __device__ inline long shell_count(const Blk &b) {
  Shell s = make_shell(b);
  return s.na + s.nb + s.nc;
}

__host__ __device__ inline long shell_entry(const Blk &b, const Shell &s, long t) {
  long lk, lj, li;
  if (t < s.na) {
    long p = t / s.nijd;
    lk = p < s.lk0 ? p : s.lk1 + (p - s.lk0);
    return b.npc + lk * s.nijd + t % s.nijd;
  }
  t -= s.na;
  if (t < s.nb) {
    long rb = s.lj0 + (s.njd - s.lj1), row = t / s.nid, r = row % rb;
    lk = s.lk0 + row / rb;
    lj = r < s.lj0 ? r : s.lj1 + (r - s.lj0);
    li = t % s.nid;
  } else {
    t -= s.nb;
    long cc = s.li0 + (s.nid - s.li1), q = t / cc, c = t % cc, nj = s.lj1 - s.lj0;
    lj = s.lj0 + q % nj;
    lk = s.lk0 + q / nj;
    li = c < s.li0 ? c : s.li1 + (c - s.li0);
  }
  return b.npc + lk * s.nijd + lj * s.nid + li;
}

// k planes per workgroup of a k-marching kernel: kmax, halved until the
// grid of tiles_xy i-j tiles has at least 512 workgroups (two per compute
// unit of MI355X's 256) or the slab is one plane.  A deeper slab re-reads
// fewer planes below its first; a small block needs more workgroups to fill
// the chip.  The coefficient and sub-iteration kernels form a slab's first
// plane with the functions that form the planes they carry, so their values
// do not depend on the depth (their all-blocks launches keep full slabs).
// The fused residual forms the k face below a slab in other code than the
// faces it carries, which may round it differently (a few ulp): its
// all-blocks launch gives each block the per-block depth.
inline int slab_depth(long tiles_xy, int nz, int kmax) {
  int ks = kmax;
  while (ks > 1 && tiles_xy * ((nz + ks - 1) / ks) < 512) ks /= 2;
  return ks;
}

// Several blocks in one flattened 1-D launch per pass.  A device table holds
// one entry per block (type E: members start[pass], gx[pass], gy[pass]): the
// block's first workgroup in the launch and its own grid (x, y; z follows).
// Linear workgroup index, remapped so that each XCD owns a contiguous run
// of the launch (neighbouring tiles of a block share its L2).
__device__ __forceinline__ unsigned xcd_linear() {
  unsigned nwg = gridDim.x, hw = blockIdx.x, x = hw % 8u, q = nwg / 8u, r = nwg % 8u;
  return x * q + (x < r ? x : r) + hw / 8u;
}

// entry of the block that owns workgroup wg of a pass (bisection over start)
template <typename E>
__device__ __forceinline__ int tab_find(const E *__restrict__ tab, int nb, int pass, unsigned wg) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tab[mid].start[pass] <= wg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// workgroup wg's tile within its block's own grid
template <typename E>
__device__ __forceinline__ dim3 tab_tile(const E &e, int pass, unsigned wg) {
  unsigned loc = wg - e.start[pass], gx = e.gx[pass], gy = e.gy[pass];
  return dim3(loc % gx, (loc / gx) % gy, loc / (gx * gy));
}

}  // namespace codens

// p[e ld + n] = 0 for the ghost entries n of block b within n0..n1, planes
// e < np (abi.hip)
__global__ void k_zero_shell(codens::Blk b, long n0, long n1, double *p, long ld, int np);
void codens_zero_shell(const codens::Blk &b, long n0, long n1, double *p, long ld, int np, hipStream_t st);

// error reporting shared by the entry points (abi.cpp)
extern "C" void codens_set_error(const char *fmt, ...);
int codens_check_launch(const char *what);
int codens_check_block(const codens_block *b, const char *what);
void *codens_workspace(size_t bytes, int slot);
// Device copy of a per-layout table, made on the first call for a key and
// kept for the process on the current device; nullptr (with the error set)
// if the upload fails or the first call for a key comes while st is captured.
void *codens_device_table(int kind, const std::vector<int32_t> &key, const void *data, size_t bytes,
                          hipStream_t st, const char *what);
bool codens_timing_on();
// Two device counters of a persistent launch's tile queue, per device and
// workspace set, zeroed when made (on the first call, which may not come
// while st is captured); the launch leaves them at zero again.
unsigned *codens_queue_counters(hipStream_t st, const char *what);
// workgroups resident on the current device at per_cu per compute unit
int codens_resident_workgroups(int per_cu);
// implimf_3d with the caller's scratch of 9 nnn doubles (nothing allocated)
int codens_implimf_3d_with(const codens_block *b, const codens_phys *ph, const double *u, const double *dt, double *v,
                           int64_t ldv, double *d, const double *ff, const double *mu, const double *mut, int32_t lmx,
                           const double *sn_blk, const double *vol, double dtpas, int32_t kfmg, int32_t ityprk,
                           const double *cson, void *stream, double *ws, const char *what);
void *codens_timing_begin(const char *name, hipStream_t st);
void codens_timing_end(const char *name, hipStream_t st);

// Brackets one kernel launch with timing events when timing is enabled.
struct CodensTimed {
  const char *name;
  hipStream_t st;
  CodensTimed(const char *n, hipStream_t s) : name(n), st(s) {
    if (n) codens_timing_begin(n, s);
  }
  ~CodensTimed() {
    if (name) codens_timing_end(name, st);
  }
};
