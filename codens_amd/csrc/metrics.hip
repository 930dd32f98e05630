// Mesh metrics of one block on the device: the zero-thickness ghost nodes
// (extmhg.f90:88-125 with ex1 = 1, ex2 = 0: copies of the nearest real
// node), the face area vectors sn (norm.f90:102-286) and the cell volumes
// (svol.f90:90-149).  One thread per storage node or cell.  Every value is
// formed with the reference's operations in its order, without contraction,
// so a mesh moved on the device (the adjoint's wall displacements) gets the
// metrics the host restatement gives, bit for bit.
#include "common.hpp"

using namespace codens;

namespace {

struct Nodes {
  double *x, *y, *z;
};

// extmhg.f90:88-125, pass 0: the k ghost layers (k1-1 from k1, k2+1 from
// k2) over the real nodes' (i, j); pass 1: the i ghost layers over j real,
// every k; pass 2: the j ghost layers over every i and k.  blockIdx.y picks
// the low or the high layer; x runs over the layer's nodes.
__global__ void __launch_bounds__(256) k_ghost_nodes(Blk b, Nodes P, int pass) {
  // node extents: real nodes i1..i2 (cells i1..i2-1), ghosts i1-1 and i2+1
  const int hi = blockIdx.y;
  long t = blockIdx.x * 256L + threadIdx.x;
  int i, j, k;
  long src;
  if (pass == 0) {
    int ni = b.i2 - b.i1 + 1, nj = b.j2 - b.j1 + 1;
    if (t >= (long)ni * nj) return;
    i = b.i1 + (int)(t % ni); j = b.j1 + (int)(t / ni);
    k = hi ? b.k2 + 1 : b.k1 - 1;
    src = b.idx(i, j, hi ? b.k2 : b.k1);
  } else if (pass == 1) {
    int nj = b.j2 - b.j1 + 1, nk = b.k2 - b.k1 + 3;
    if (t >= (long)nj * nk) return;
    j = b.j1 + (int)(t % nj); k = b.k1 - 1 + (int)(t / nj);
    i = hi ? b.i2 + 1 : b.i1 - 1;
    src = b.idx(hi ? b.i2 : b.i1, j, k);
  } else {
    int ni = b.i2 - b.i1 + 3, nk = b.k2 - b.k1 + 3;
    if (t >= (long)ni * nk) return;
    i = b.i1 - 1 + (int)(t % ni); k = b.k1 - 1 + (int)(t / ni);
    j = hi ? b.j2 + 1 : b.j1 - 1;
    src = b.idx(i, hi ? b.j2 : b.j1, k);
  }
  long n = b.idx(i, j, k);
  P.x[n] = P.x[src];
  P.y[n] = P.y[src];
  P.z[n] = P.z[src];
}

struct V3 {
  double x, y, z;
};

__device__ inline V3 at(const Nodes &P, long n) { return V3{P.x[n], P.y[n], P.z[n]}; }
__device__ inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 cross(V3 d1, V3 d2) {
#pragma clang fp contract(off)
  return V3{d1.y * d2.z - d1.z * d2.y, d1.z * d2.x - d1.x * d2.z, d1.x * d2.y - d1.y * d2.x};
}

// norm.f90's orientation factor of one face family: +-0.5 by the sign of
// the family's normal at the test node (i1+1, j1+1, k1+1) against the sum of
// the node differences along the third direction (norm.f90:133-156; the
// other families likewise).  The four differences are summed pairwise in
// order, as the host restatement does; only the sign is kept.
__device__ inline double orient(const Nodes &P, long o, long a1, long a2, long a3, const long e[4][2]) {
#pragma clang fp contract(off)
  V3 t = cross(sub(at(P, o + a1), at(P, o)), sub(at(P, o + a2), at(P, o + a3)));
  V3 w = sub(at(P, o + e[0][0]), at(P, o + e[0][1]));
  for (int q = 1; q < 4; q++) {
    V3 d = sub(at(P, o + e[q][0]), at(P, o + e[q][1]));
    w = V3{w.x + d.x, w.y + d.y, w.z + d.z};
  }
  double ps = t.x * w.x + t.y * w.y + t.z * w.z;
  return ps >= 0. ? 0.5 : -0.5;
}

// sn of the storage node (i, j, k), local offsets li, lj, lk: the k face
// (norm.f90:158-176), the j face (:205-223) and the i face (:252-270) whose
// low corner is the node, zero where the face leaves the storage.
__global__ void __launch_bounds__(256) k_metric_sn(Blk b, Nodes P, double *__restrict__ sn) {
#pragma clang fp contract(off)
  long m = blockIdx.x * 256L + threadIdx.x;
  if (m >= b.nnn) return;
  const long nid = b.nid, nijd = b.nijd, njd = nijd / nid, nkd = b.nnn / nijd;
  const int li = (int)(m % nid), lj = (int)((m / nid) % njd), lk = (int)(m / nijd);
  const long n = b.npc + m;
  const long I = 1, J = nid, K = nijd;
  const long o = b.idx(b.i1 + 1, b.j1 + 1, b.k1 + 1);
  double s[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};  // [direction][component]
  if (lj < njd - 1 && li < nid - 1) {  // k face
    const long e[4][2] = {{K, 0}, {I + K, I}, {J + K, J}, {I + J + K, I + J}};
    double sg = orient(P, o, I + J, J, I, e);
    V3 t = cross(sub(at(P, n + I + J), at(P, n)), sub(at(P, n + J), at(P, n + I)));
    s[2][0] = sg * t.x; s[2][1] = sg * t.y; s[2][2] = sg * t.z;
  }
  if (lk < nkd - 1 && li < nid - 1) {  // j face
    const long e[4][2] = {{J, 0}, {I + J, I}, {J + K, K}, {I + J + K, I + K}};
    double sg = orient(P, o, I + K, K, I, e);
    V3 t = cross(sub(at(P, n + I + K), at(P, n)), sub(at(P, n + K), at(P, n + I)));
    s[1][0] = sg * t.x; s[1][1] = sg * t.y; s[1][2] = sg * t.z;
  }
  if (lk < nkd - 1 && lj < njd - 1) {  // i face
    const long e[4][2] = {{I, 0}, {I + J, J}, {I + K, K}, {I + J + K, J + K}};
    double sg = orient(P, o, J + K, K, J, e);
    V3 t = cross(sub(at(P, n + J + K), at(P, n)), sub(at(P, n + K), at(P, n + J)));
    s[0][0] = sg * t.x; s[0][1] = sg * t.y; s[0][2] = sg * t.z;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll
    for (int d = 0; d < 3; d++) sn[((long)c * 3 + d) * b.nnn + m] = s[d][c];
  }
}

// svol.f90:90-136: (sum of the face's four corners) . S, the corners added
// in order, the components x, y, z in order
__device__ inline double face_sum(const Nodes &P, const double *__restrict__ sn, long nnn, long n, long m, int d,
                                  long c1, long c2, long c3) {
#pragma clang fp contract(off)
  const double *A[3] = {P.x, P.y, P.z};
  double acc = 0.;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double *a = A[c];
    double s = ((a[n] + a[n + c1]) + a[n + c2]) + a[n + c3];
    double term = s * sn[((long)c * 3 + d) * nnn + m];
    acc = c == 0 ? term : acc + term;
  }
  return acc;
}

// svol.f90:138-149: vol = 0.25 (siv(i+1) - siv + sjv(j+1) - sjv + skv(k+1)
// - skv) / 3 over the cells k1-1..k2, j1-1..j2, i1-1..i2 (every storage cell
// but the last layer of each direction, which gets 0)
__global__ void __launch_bounds__(256) k_metric_vol(Blk b, Nodes P, const double *__restrict__ sn,
                                                    double *__restrict__ vol) {
#pragma clang fp contract(off)
  long m = blockIdx.x * 256L + threadIdx.x;
  if (m >= b.nnn) return;
  const long nid = b.nid, nijd = b.nijd, njd = nijd / nid, nkd = b.nnn / nijd;
  const int li = (int)(m % nid), lj = (int)((m / nid) % njd), lk = (int)(m / nijd);
  const long n = b.npc + m;
  if (li >= nid - 1 || lj >= njd - 1 || lk >= nkd - 1) {
    vol[n] = 0.;
    return;
  }
  const long I = 1, J = nid, K = nijd;
  double si1 = face_sum(P, sn, b.nnn, n, m, 0, J, K, J + K);
  double si2 = face_sum(P, sn, b.nnn, n + I, m + I, 0, J, K, J + K);
  double sj1 = face_sum(P, sn, b.nnn, n, m, 1, I, K, I + K);
  double sj2 = face_sum(P, sn, b.nnn, n + J, m + J, 1, I, K, I + K);
  double sk1 = face_sum(P, sn, b.nnn, n, m, 2, I, J, I + J);
  double sk2 = face_sum(P, sn, b.nnn, n + K, m + K, 2, I, J, I + J);
  vol[n] = 0.25 * (si2 - si1 + sj2 - sj1 + sk2 - sk1) / 3;
}

}  // namespace

extern "C" int codens_mesh_metrics(const codens_block *cb, double *x, double *y, double *z, double *sn_blk,
                                   double *vol, void *stream) {
  int rc = codens_check_block(cb, "codens_mesh_metrics");
  if (rc) return rc;
  if (!x || !y || !z || !sn_blk || !vol) {
    codens_set_error("codens_mesh_metrics: null argument");
    return CODENS_EINVAL;
  }
  Blk b = make_blk(cb);
  if (b.i1 - 1 < b.id1 || b.j1 - 1 < b.jd1 || b.k1 - 1 < b.kd1 || b.i2 + 1 > b.id1 + (int)b.nid - 1 ||
      b.j2 + 1 > b.jd1 + (int)(b.nijd / b.nid) - 1 || b.k2 + 1 > b.kd1 + (int)(b.nnn / b.nijd) - 1 ||
      b.i2 - b.i1 < 1 || b.j2 - b.j1 < 1 || b.k2 - b.k1 < 1) {
    codens_set_error("codens_mesh_metrics: the block needs one ghost node layer on every side and at least one "
                     "cell per direction");
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  Nodes P = {x, y, z};
  long nl[3] = {(long)(b.i2 - b.i1 + 1) * (b.j2 - b.j1 + 1), (long)(b.j2 - b.j1 + 1) * (b.k2 - b.k1 + 3),
                (long)(b.i2 - b.i1 + 3) * (b.k2 - b.k1 + 3)};
  for (int pass = 0; pass < 3; pass++)
    hipLaunchKernelGGL(k_ghost_nodes, dim3((unsigned)((nl[pass] + 255) / 256), 2), dim3(256), 0, st, b, P, pass);
  dim3 g((unsigned)((b.nnn + 255) / 256));
  hipLaunchKernelGGL(k_metric_sn, g, dim3(256), 0, st, b, P, sn_blk);
  hipLaunchKernelGGL(k_metric_vol, g, dim3(256), 0, st, b, P, (const double *)sn_blk, vol);
  return codens_check_launch("codens_mesh_metrics");
}
