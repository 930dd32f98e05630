// codens_implimf_3d_noalloc: the implicit phase of codens_implimf_3d with no
// allocation, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).  Its scratch --
// coefficients (3), diagonal (1) and the ping-pong increment (5) of the
// block, 9 doubles per storage entry -- is a file-scope device array sized
// for blocks of up to NOALLOC_MAX_NNN storage entries (9 x 8 x 2,400,000 B =
// 172.8 MB on each device the library runs on, held until the process ends).
//
// Every call on a device shares that device's array: calls that overlap on
// one device (concurrent streams, or two graphs replayed at once) race.  The
// array keeps the last call's contents; nothing in it has to start at zero:
// the coefficient pass writes every coefficient the sub-iterations read, and
// the sub-iterations read increments of real cells only, each written by the
// sub-iteration before.
#include <mutex>

#include "common.hpp"

constexpr long NOALLOC_MAX_NNN = 2400000;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[9 * NOALLOC_MAX_NNN];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES];  // the device address of `scratch`, per device
}  // namespace

// Device address of the scratch array on the current device, looked up once
// per device with that device current (each device holds its own copy).
static double *scratch_on_current_device(const char *what) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    codens_set_error("%s: no usable current device", what);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      codens_set_error("%s: hipGetSymbolAddress of the scratch array failed on device %d: %s", what, dev,
                       hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = (double *)p;
  }
  return g_scratch_ptr[dev];
}

extern "C" int codens_implimf_3d_noalloc(const codens_block *b, const codens_phys *ph, const double *u,
                                         const double *dt, double *v, int64_t ldv, double *d, const double *ff,
                                         const double *mu, const double *mut, int32_t lmx, const double *sn_blk,
                                         const double *vol, double dtpas, int32_t kfmg, int32_t ityprk,
                                         const double *ps, const double *cson, void *stream) {
  const char *what = "codens_implimf_3d_noalloc";
  (void)ps;  // as codens_implimf_3d: p of the real cells from v
  int rc = codens_check_block(b, what);
  if (rc) return rc;
  if ((long)b->nnn > NOALLOC_MAX_NNN) {
    codens_set_error("%s: block of %d storage entries exceeds the scratch's maximum of %ld (9 doubles each); "
                     "use codens_implimf_3d", what, b->nnn, NOALLOC_MAX_NNN);
    return CODENS_EINVAL;
  }
  double *ws = scratch_on_current_device(what);
  if (!ws) return CODENS_EHIP;
  return codens_implimf_3d_with(b, ph, u, dt, v, ldv, d, ff, mu, mut, lmx, sn_blk, vol, dtpas, kfmg, ityprk, cson,
                                stream, ws, what);
}
