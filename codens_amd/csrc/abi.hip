// Error reporting, device workspace and the reductions of the C-ABI.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "common.hpp"

using namespace codens;

static thread_local char g_err[512] = "";

extern "C" void codens_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char *codens_last_error(void) { return g_err; }

int codens_check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    codens_set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

int codens_check_block(const codens_block *b, const char *what) {
  if (!b) {
    codens_set_error("%s: null block", what);
    return CODENS_EINVAL;
  }
  long nid = b->id2 - b->id1 + 1, njd = b->jd2 - b->jd1 + 1, nkd = b->kd2 - b->kd1 + 1;
  if (b->npc < 0 || nid < 4 || njd < 4 || nkd < 4 || (long)b->nnn != nid * njd * nkd ||
      b->ii1 - b->id1 < 1 || b->id2 - b->ii2 < 1 || b->jj1 - b->jd1 < 1 || b->jd2 - b->jj2 < 1 ||
      b->kk1 - b->kd1 < 1 || b->kd2 - b->kk2 < 1 || b->ii2 - b->ii1 < 2 || b->jj2 - b->jj1 < 2 ||
      b->kk2 - b->kk1 < 2) {
    codens_set_error("%s: inconsistent block extents (npc=%d nnn=%d)", what, b->npc, b->nnn);
    return CODENS_EINVAL;
  }
  return CODENS_OK;
}

// Grow-only device workspaces, kept until process exit: per device, up to
// CODENS_WS_SETS independent sets of slots.  The entry points draw from the
// set selected by codens_workspace_set (0 by default, per host thread); a
// caller that runs blocks concurrently on several streams gives each stream
// its own set.
//
// A slot that must grow gets a new buffer and the old one is retired, not
// freed: a HIP graph captured earlier holds the old address, and its replays
// must keep writing memory that nothing else is given.  So growing never
// synchronises or frees, and is also safe while another stream runs.
namespace {
constexpr int CODENS_WS_SETS = 16;
struct Slot {
  void *p = nullptr;
  size_t n = 0;
};
std::mutex g_ws_mu;
std::vector<std::vector<std::vector<Slot>>> g_ws;  // [device][set][slot]
std::vector<void *> g_ws_retired;                  // outgrown buffers, live to the end
thread_local int g_ws_set = 0;
}  // namespace

extern "C" int codens_workspace_set(int32_t set) {
  if (set < 0 || set >= CODENS_WS_SETS) {
    codens_set_error("codens_workspace_set: set %d is outside 0..%d", set, CODENS_WS_SETS - 1);
    return CODENS_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_ws_mu);
  g_ws_set = set;
  return CODENS_OK;
}

void *codens_workspace(size_t bytes, int slot) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
  if ((int)g_ws[dev].size() <= g_ws_set) g_ws[dev].resize(g_ws_set + 1);
  auto &v = g_ws[dev][g_ws_set];
  if ((int)v.size() <= slot) v.resize(slot + 1);
  Slot &s = v[slot];
  if (s.n < bytes) {
    void *np_ = nullptr;
    if (hipMalloc(&np_, bytes) != hipSuccess) return nullptr;
    if (s.p) g_ws_retired.push_back(s.p);
    s.p = np_;
    s.n = bytes;
  }
  return s.p;
}

__global__ void __launch_bounds__(256) k_zero_shell(codens::Blk b, long n0, long n1, double *p, long ld, int np) {
  const codens::Shell s = codens::make_shell(b);
  long t = blockIdx.x * 256L + threadIdx.x;
  if (t >= s.na + s.nb + s.nc) return;
  long n = codens::shell_entry(b, s, t);
  if (n < n0 || n > n1) return;
  for (int e = 0; e < np; e++) p[e * ld + n] = 0.;
}

void codens_zero_shell(const codens::Blk &b, long n0, long n1, double *p, long ld, int np, hipStream_t st) {
  long cnt = codens::shell_count(b);
  if (cnt > 0)
    hipLaunchKernelGGL(k_zero_shell, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, b, n0, n1, p, ld, np);
}

namespace {
std::map<std::pair<int, int>, unsigned *> g_queue;  // (device, workspace set) -> 2 counters
}  // namespace

unsigned *codens_queue_counters(hipStream_t st, const char *what) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    codens_set_error("%s: no current device", what);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto k = std::make_pair(dev, g_ws_set);
  auto it = g_queue.find(k);
  if (it != g_queue.end()) return it->second;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    codens_set_error("%s: first persistent launch while the stream is captured; call it once before", what);
    return nullptr;
  }
  unsigned *p = nullptr;
  if (hipMalloc(&p, 2 * sizeof(unsigned)) != hipSuccess || hipMemset(p, 0, 2 * sizeof(unsigned)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    if (p) (void)hipFree(p);
    codens_set_error("%s: tile queue allocation failed", what);
    return nullptr;
  }
  g_queue[k] = p;
  return p;
}

int codens_resident_workgroups(int per_cu) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    ncu = 256;
  return per_cu * ncu;
}

namespace {
std::mutex g_tab_mu;
// device tables by (device, kind, key); made once, kept to the end
std::map<std::tuple<int, int, std::vector<int32_t>>, void *> g_tab;
}  // namespace

void *codens_device_table(int kind, const std::vector<int32_t> &key, const void *data, size_t bytes,
                          hipStream_t st, const char *what) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    codens_set_error("%s: no current device", what);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto k = std::make_tuple(dev, kind, key);
  auto it = g_tab.find(k);
  if (it != g_tab.end()) return it->second;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    codens_set_error("%s: first call for these blocks while the stream is captured; call it once before", what);
    return nullptr;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess || hipMemcpy(p, data, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    if (p) (void)hipFree(p);
    codens_set_error("%s: block table upload failed", what);
    return nullptr;
  }
  g_tab[k] = p;
  return p;
}

// ---------------------------------------------------------------------------
// Reductions over the real cells: one partial per workgroup, summed on the
// host in workgroup order so the result does not depend on timing.
constexpr int RED_T = 256;

template <int OP>  // 0: min of a, 1: sum of ((a-b)/dt)^2
__global__ void __launch_bounds__(RED_T) k_reduce(Blk b, const double *__restrict__ a,
                                                  const double *__restrict__ bb,
                                                  const double *__restrict__ dt, double *partial) {
  __shared__ double sh[RED_T / 64];
  long nx = b.i2 - b.i1, ny = b.j2 - b.j1, nz = b.k2 - b.k1;
  long total = nx * ny * nz;
  double acc = OP == 0 ? DBL_MAX : 0.0;
  for (long t = blockIdx.x * (long)RED_T + threadIdx.x; t < total; t += (long)gridDim.x * RED_T) {
    int i = b.i1 + (int)(t % nx);
    long r = t / nx;
    int j = b.j1 + (int)(r % ny);
    int k = b.k1 + (int)(r / ny);
    long n = b.idx(i, j, k);
    if (OP == 0) {
      acc = fmin(acc, a[n]);
    } else {
      double x = (a[n] - bb[n]) / dt[n];
      acc += x * x;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(acc, off, 64);
    acc = OP == 0 ? fmin(acc, o) : acc + o;
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = sh[0];
    for (int q = 1; q < RED_T / 64; q++) s = OP == 0 ? fmin(s, sh[q]) : s + sh[q];
    partial[blockIdx.x] = s;
  }
}

template <int OP>
static int reduce_blocks(const codens_block *blocks, int nb, const double *a, const double *bb,
                         const double *dt, double *out, hipStream_t st, const char *what) {
  const int G = 512;
  double *part = (double *)codens_workspace(sizeof(double) * G * nb, 15);
  if (!part) {
    codens_set_error("%s: workspace allocation failed", what);
    return CODENS_EHIP;
  }
  for (int l = 0; l < nb; l++) {
    int rc = codens_check_block(&blocks[l], what);
    if (rc) return rc;
    Blk b = make_blk(&blocks[l]);
    hipLaunchKernelGGL(k_reduce<OP>, dim3(G), dim3(RED_T), 0, st, b, a, bb, dt, part + (long)l * G);
    if ((rc = codens_check_launch(what))) return rc;
  }
  std::vector<double> h((size_t)G * nb);
  if (hipMemcpyAsync(h.data(), part, sizeof(double) * G * nb, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    codens_set_error("%s: copy of partial results failed", what);
    return CODENS_EHIP;
  }
  double s = OP == 0 ? DBL_MAX : 0.0;
  for (double x : h) s = OP == 0 ? fmin(s, x) : s + x;
  *out = s;
  return CODENS_OK;
}

__global__ void k_fill(double *p, long n0, long n1, double val) {
  long n = n0 + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n <= n1) p[n] = val;
}

extern "C" int codens_chrono_global(const codens_block *blocks, int32_t nb, double dt1min, double *dt,
                                    double *dtmin_out, void *stream) {
  if (!blocks || nb <= 0 || !dt) {
    codens_set_error("codens_chrono_global: invalid argument");
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  double dtmin = 0;
  int rc = reduce_blocks<0>(blocks, nb, dt, nullptr, nullptr, &dtmin, st, "codens_chrono_global");
  if (rc) return rc;
  if (dtmin_out) *dtmin_out = dtmin;
  double dtm = fmin(dt1min, dtmin);
  if (dtm < dt1min) {
    codens_set_error("codens_chrono_global: time step %.6e below dt1min %.6e", dtmin, dt1min);
    return CODENS_ESTOP;
  }
  for (int l = 0; l < nb; l++) {
    Blk b = make_blk(&blocks[l]);
    long n0 = b.npc, n1 = b.npc + b.nnn - 1;
    long cnt = n1 - n0 + 1;
    hipLaunchKernelGGL(k_fill, dim3((cnt + 255) / 256), dim3(256), 0, st, dt, n0, n1, dtm);
    if ((rc = codens_check_launch("codens_chrono_global"))) return rc;
  }
  return CODENS_OK;
}

extern "C" int codens_sumsq_rate(const codens_block *blocks, int32_t nb, const double *a, const double *bv,
                                 const double *dt, double *out, void *stream) {
  if (!blocks || nb <= 0 || !a || !bv || !dt || !out) {
    codens_set_error("codens_sumsq_rate: invalid argument");
    return CODENS_EINVAL;
  }
  return reduce_blocks<1>(blocks, nb, a, bv, dt, out, (hipStream_t)stream, "codens_sumsq_rate");
}

// ---------------------------------------------------------------------------
// Per-kernel launch timing with HIP events recorded on the launch stream.
namespace {
struct Timed {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};
std::mutex g_tm_mu;
bool g_tm_on = false;
std::map<std::string, Timed> g_tm;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_tm_pool;
}  // namespace

bool codens_timing_on() { return g_tm_on; }

void *codens_timing_begin(const char *name, hipStream_t st) {
  if (!g_tm_on) return nullptr;
  std::lock_guard<std::mutex> lk(g_tm_mu);
  std::pair<hipEvent_t, hipEvent_t> p;
  if (!g_tm_pool.empty()) {
    p = g_tm_pool.back();
    g_tm_pool.pop_back();
  } else {
    if (hipEventCreate(&p.first) != hipSuccess || hipEventCreate(&p.second) != hipSuccess) return nullptr;
  }
  auto &t = g_tm[name];
  t.ev.push_back(p);
  (void)hipEventRecord(p.first, st);
  return (void *)&t.ev.back();
}

void codens_timing_end(const char *name, hipStream_t st) {
  if (!g_tm_on) return;
  std::lock_guard<std::mutex> lk(g_tm_mu);
  auto it = g_tm.find(name);
  if (it == g_tm.end() || it->second.ev.empty()) return;
  (void)hipEventRecord(it->second.ev.back().second, st);
}

extern "C" int codens_timing_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_tm_mu);
  g_tm_on = on != 0;
  return CODENS_OK;
}

extern "C" int codens_timing_read(const char *name, double *ms_total, int64_t *count, int32_t reset) {
  if (!name || !ms_total || !count) {
    codens_set_error("codens_timing_read: null argument");
    return CODENS_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_tm_mu);
  *ms_total = 0.0;
  *count = 0;
  auto it = g_tm.find(name);
  if (it == g_tm.end()) return CODENS_OK;
  for (auto &p : it->second.ev) {
    if (hipEventSynchronize(p.second) != hipSuccess) {
      codens_set_error("codens_timing_read: event synchronisation failed");
      return CODENS_EHIP;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.first, p.second) != hipSuccess) {
      codens_set_error("codens_timing_read: elapsed time failed");
      return CODENS_EHIP;
    }
    *ms_total += ms;
    *count += 1;
  }
  if (reset) {
    for (auto &p : it->second.ev) g_tm_pool.push_back(p);
    it->second.ev.clear();
  }
  return CODENS_OK;
}
