// On-disk formats of the reference: restart records (writda.f90 / readda.f90)
// and mesh records (writdg.f90 / readdg.f90).  Host code only: the arrays are
// host memory in the reference storage; a device-resident driver copies its
// state out before writing and in after reading.
//
// The files are Fortran unformatted sequential: every record is its length
// in bytes as a 4-byte integer, the data, and the length again (the layout
// gfortran and flang write on little-endian hosts).  A record holds one
// variable of one block over a box of cells or nodes, i fastest, then j, k
// (the implied-DO order of writda.f90:46-66 and writdg.f90:44-46).
#include <stdio.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "common.hpp"

using namespace codens;

namespace {

struct Box {
  int i1, i2, j1, j2, k1, k2;  // inclusive
  long count() const { return (long)(i2 - i1 + 1) * (j2 - j1 + 1) * (k2 - k1 + 1); }
};

// real cells ii1..ii2-1 (chdacc.f90 'cccc', readda.f90:43-58)
Box cell_box(const codens_block &b) { return {b.ii1, b.ii2 - 1, b.jj1, b.jj2 - 1, b.kk1, b.kk2 - 1}; }
// nodes ii1..ii2 (readdg.f90:58-66)
Box node_box(const codens_block &b) { return {b.ii1, b.ii2, b.jj1, b.jj2, b.kk1, b.kk2}; }

long entry(const codens_block &b, int i, int j, int k) {
  long nid = b.id2 - b.id1 + 1, nijd = nid * (b.jd2 - b.jd1 + 1);
  return b.npc + (i - b.id1) + (j - b.jd1) * nid + (k - b.kd1) * nijd;
}

int write_record(FILE *f, const codens_block &b, const Box &x, const double *a, std::vector<double> &buf,
                 const char *what, const char *path) {
  long n = x.count();
  if (n * 8 > INT32_MAX) {
    codens_set_error("%s: record of %ld values exceeds the 2 GiB of one record", what, n);
    return CODENS_EINVAL;
  }
  buf.resize(n);
  long p = 0;
  for (int k = x.k1; k <= x.k2; k++)
    for (int j = x.j1; j <= x.j2; j++) {
      long e = entry(b, x.i1, j, k);
      for (int i = x.i1; i <= x.i2; i++) buf[p++] = a[e++];
    }
  int32_t len = (int32_t)(n * 8);
  if (fwrite(&len, 4, 1, f) != 1 || fwrite(buf.data(), 8, n, f) != (size_t)n || fwrite(&len, 4, 1, f) != 1) {
    codens_set_error("%s: write to %s failed", what, path);
    return CODENS_EIO;
  }
  return CODENS_OK;
}

int read_record(FILE *f, const codens_block &b, const Box &x, double *a, std::vector<double> &buf, const char *what,
                const char *path, int l, const char *name) {
  long n = x.count();
  int32_t len = 0, len2 = 0;
  if (fread(&len, 4, 1, f) != 1) {
    codens_set_error("%s: %s ends before record %s of block %d", what, path, name, l + 1);
    return CODENS_EIO;
  }
  if ((long)len != n * 8) {
    codens_set_error("%s: record %s of block %d in %s holds %d bytes, the block needs %ld", what, name, l + 1, path,
                     len, n * 8);
    return CODENS_EIO;
  }
  buf.resize(n);
  if (fread(buf.data(), 8, n, f) != (size_t)n || fread(&len2, 4, 1, f) != 1 || len2 != len) {
    codens_set_error("%s: record %s of block %d in %s is truncated or its end marker is wrong", what, name, l + 1,
                     path);
    return CODENS_EIO;
  }
  long p = 0;
  for (int k = x.k1; k <= x.k2; k++)
    for (int j = x.j1; j <= x.j2; j++) {
      long e = entry(b, x.i1, j, k);
      for (int i = x.i1; i <= x.i2; i++) a[e++] = buf[p++];
    }
  return CODENS_OK;
}

int check_blocks(const codens_block *blocks, int32_t nb, int64_t ld, const char *what) {
  if (!blocks || nb <= 0 || ld <= 0) {
    codens_set_error("%s: invalid block list", what);
    return CODENS_EINVAL;
  }
  for (int l = 0; l < nb; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    if ((int64_t)blocks[l].npc + blocks[l].nnn > ld) {
      codens_set_error("%s: block %d lies beyond the arrays (npc %d + nnn %d > %lld)", what, l + 1, blocks[l].npc,
                       blocks[l].nnn, (long long)ld);
      return CODENS_EINVAL;
    }
  }
  return CODENS_OK;
}

}  // namespace

extern "C" int codens_restart_write(const char *path, const codens_block *blocks, int32_t nb, int32_t ns,
                                    int32_t nturb, const double *v, int64_t ldv, const double *mut) {
  const char *what = "codens_restart_write";
  int rc = check_blocks(blocks, nb, ldv, what);
  if (rc) return rc;
  if (!path || !v || (ns && !mut) || (nturb != 0 && nturb != 2)) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  FILE *f = fopen(path, "wb");  // the reference rewinds for its first block (writda.f90:28)
  if (!f) {
    codens_set_error("%s: cannot open %s for writing", what, path);
    return CODENS_EIO;
  }
  std::vector<double> buf;
  for (int l = 0; l < nb && !rc; l++) {
    const codens_block &b = blocks[l];
    Box x = cell_box(b);
    for (int e = 0; e < 5 && !rc; e++) rc = write_record(f, b, x, v + e * ldv, buf, what, path);
    if (!rc && ns) rc = write_record(f, b, x, mut, buf, what, path);
    for (int e = 5; e < 5 + nturb && !rc; e++) rc = write_record(f, b, x, v + e * ldv, buf, what, path);
  }
  if (fclose(f) != 0 && !rc) {
    codens_set_error("%s: closing %s failed", what, path);
    rc = CODENS_EIO;
  }
  return rc;
}

extern "C" int codens_restart_read(const char *path, const codens_block *blocks, int32_t nb, int32_t ns,
                                   int32_t nturb, double *v, int64_t ldv, double *mut) {
  const char *what = "codens_restart_read";
  int rc = check_blocks(blocks, nb, ldv, what);
  if (rc) return rc;
  if (!path || !v || (ns && !mut) || (nturb != 0 && nturb != 2)) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    codens_set_error("%s: cannot open %s", what, path);
    return CODENS_EIO;
  }
  static const char *names[] = {"rho", "rho u", "rho v", "rho w", "rho E", "v6", "v7"};
  std::vector<double> buf;
  for (int l = 0; l < nb && !rc; l++) {
    const codens_block &b = blocks[l];
    Box x = cell_box(b);
    for (int e = 0; e < 5 && !rc; e++) rc = read_record(f, b, x, v + e * ldv, buf, what, path, l, names[e]);
    if (!rc && ns) rc = read_record(f, b, x, mut, buf, what, path, l, "mut");
    for (int e = 5; e < 5 + nturb && !rc; e++) rc = read_record(f, b, x, v + e * ldv, buf, what, path, l, names[e]);
  }
  fclose(f);
  return rc;
}

extern "C" int codens_mesh_write(const char *path, const codens_block *blocks, int32_t nb, int64_t ld,
                                 const double *x, const double *y, const double *z) {
  const char *what = "codens_mesh_write";
  int rc = check_blocks(blocks, nb, ld, what);
  if (rc) return rc;
  if (!path || !x || !y || !z) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    codens_set_error("%s: cannot open %s for writing", what, path);
    return CODENS_EIO;
  }
  std::vector<double> buf;
  for (int l = 0; l < nb && !rc; l++) {
    Box bx = node_box(blocks[l]);
    const double *c[3] = {x, y, z};
    for (int q = 0; q < 3 && !rc; q++) rc = write_record(f, blocks[l], bx, c[q], buf, what, path);
  }
  if (fclose(f) != 0 && !rc) {
    codens_set_error("%s: closing %s failed", what, path);
    rc = CODENS_EIO;
  }
  return rc;
}

extern "C" int codens_mesh_read(const char *path, const codens_block *blocks, int32_t nb, int64_t ld, double *x,
                                double *y, double *z) {
  const char *what = "codens_mesh_read";
  int rc = check_blocks(blocks, nb, ld, what);
  if (rc) return rc;
  if (!path || !x || !y || !z) {
    codens_set_error("%s: invalid argument", what);
    return CODENS_EINVAL;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    codens_set_error("%s: cannot open %s", what, path);
    return CODENS_EIO;
  }
  std::vector<double> buf;
  static const char *names[] = {"x", "y", "z"};
  for (int l = 0; l < nb && !rc; l++) {
    Box bx = node_box(blocks[l]);
    double *c[3] = {x, y, z};
    for (int q = 0; q < 3 && !rc; q++) rc = read_record(f, blocks[l], bx, c[q], buf, what, path, l, names[q]);
  }
  fclose(f);
  return rc;
}
