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
