// Host-array mirrors for callers that hold their fields in host memory (the
// Fortran drop-in routines): one device buffer per host array, found by the
// host address, grown when the caller's array is larger than before.  Only
// the window a call names is copied -- one block's slice of each variable --
// so a multi-block caller moves each block once per call, not the whole
// array.  Every transfer is one the caller asks for; the only reuse is the
// opt-in upload-once mode for arrays the caller declares constant.
#include <map>
#include <mutex>

#include "common.hpp"

namespace {
struct Mirror {
  void *dev = nullptr;
  size_t bytes = 0;
};
std::mutex g_mu;
// Keyed by (device, host address); a bundle is keyed by its first host
// array's address + 1, which no 8-byte-aligned array can start at.
std::map<std::pair<int, const void *>, Mirror> g_mirrors;

// Window of `rows` runs of `width` bytes, `pitch` apart, from byte `offset`.
struct Win {
  size_t offset, width, pitch, rows;
  size_t end() const { return offset + (rows - 1) * pitch + width; }
};

bool make_win(int64_t offset, int64_t width, int64_t pitch, int32_t rows, Win &w) {
  if (offset < 0 || width <= 0 || rows <= 0 || (rows > 1 && pitch < width)) return false;
  w = Win{(size_t)offset, (size_t)width, (size_t)(rows > 1 ? pitch : width), (size_t)rows};
  return true;
}

hipError_t copy_win(char *dst, const char *src, const Win &w, hipMemcpyKind kind, hipStream_t st) {
  if (w.rows == 1) return hipMemcpyAsync(dst + w.offset, src + w.offset, w.width, kind, st);
  return hipMemcpy2DAsync(dst + w.offset, w.pitch, src + w.offset, w.pitch, w.width, w.rows, kind, st);
}
}  // namespace

static void *mirror_of(const void *key, size_t bytes, bool *fresh = nullptr) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  Mirror &m = g_mirrors[{dev, key}];
  if (fresh) *fresh = m.bytes < bytes;
  if (m.bytes < bytes) {
    if (m.dev) {
      if (hipDeviceSynchronize() != hipSuccess) return nullptr;
      (void)hipFree(m.dev);
      m.dev = nullptr;
      m.bytes = 0;
    }
    if (hipMalloc(&m.dev, bytes) != hipSuccess) return nullptr;
    m.bytes = bytes;
  }
  return m.dev;
}

static void *find_mirror(const void *key, size_t need) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mirrors.find({dev, key});
  if (it == g_mirrors.end() || it->second.bytes < need) return nullptr;
  return it->second.dev;
}

extern "C" int codens_host_mirror(const void *host, int64_t bytes, int64_t offset, int64_t width,
                                  int64_t pitch, int32_t rows, int32_t to_device, void **dev_out,
                                  void *stream) {
  Win w;
  if (!host || bytes <= 0 || !dev_out || !make_win(offset, width, pitch, rows, w) ||
      w.end() > (size_t)bytes) {
    codens_set_error("codens_host_mirror: invalid argument or window outside the array");
    return CODENS_EINVAL;
  }
  bool fresh = false;
  char *d = (char *)mirror_of(host, (size_t)bytes, &fresh);
  if (!d) {
    codens_set_error("codens_host_mirror: device allocation of %lld bytes failed", (long long)bytes);
    return CODENS_EHIP;
  }
  hipError_t e = hipSuccess;
  if (to_device == 1)
    e = copy_win(d, (const char *)host, w, hipMemcpyHostToDevice, (hipStream_t)stream);
  else if (to_device == 2 && fresh)  // constant array: whole array, once
    e = hipMemcpyAsync(d, host, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e != hipSuccess) {
    codens_set_error("codens_host_mirror: host to device copy failed");
    return CODENS_EHIP;
  }
  *dev_out = d;
  return CODENS_OK;
}

extern "C" int codens_host_fetch(void *host, int64_t offset, int64_t width, int64_t pitch, int32_t rows,
                                 void *stream) {
  Win w;
  if (!host || !make_win(offset, width, pitch, rows, w)) {
    codens_set_error("codens_host_fetch: invalid argument");
    return CODENS_EINVAL;
  }
  const char *d = (const char *)find_mirror(host, w.end());
  if (!d) {
    codens_set_error("codens_host_fetch: no device mirror of this host array covering the window");
    return CODENS_EINVAL;
  }
  if (copy_win((char *)host, d, w, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
      hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    codens_set_error("codens_host_fetch: device to host copy failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

// Packs `count` host arrays of `bytes` each into consecutive slots of one
// device bundle (the tq and fd bundles of the C-ABI, slot stride = the
// arrays' length); the window [offset, offset+width) of each is copied.
extern "C" int codens_host_bundle(const void *const *hosts, int32_t count, int64_t bytes, int64_t offset,
                                  int64_t width, int32_t to_device, void **dev_out, void *stream) {
  if (!hosts || count <= 0 || bytes <= 0 || !dev_out || offset < 0 || width <= 0 || offset + width > bytes) {
    codens_set_error("codens_host_bundle: invalid argument or window outside the arrays");
    return CODENS_EINVAL;
  }
  for (int i = 0; i < count; i++)
    if (!hosts[i]) {
      codens_set_error("codens_host_bundle: null host array %d", i);
      return CODENS_EINVAL;
    }
  char *d = (char *)mirror_of((const char *)hosts[0] + 1, (size_t)bytes * count);
  if (!d) {
    codens_set_error("codens_host_bundle: device allocation failed");
    return CODENS_EHIP;
  }
  if (to_device)
    for (int i = 0; i < count; i++)
      if (hipMemcpyAsync(d + (size_t)i * bytes + offset, (const char *)hosts[i] + offset, (size_t)width,
                         hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess) {
        codens_set_error("codens_host_bundle: host to device copy failed");
        return CODENS_EHIP;
      }
  *dev_out = d;
  return CODENS_OK;
}

// Copies the window of every slot of a bundle made by codens_host_bundle
// back to its host array and synchronises.
extern "C" int codens_host_bundle_fetch(const void *const *hosts, int32_t count, int64_t bytes, int64_t offset,
                                        int64_t width, void *stream) {
  if (!hosts || count <= 0 || bytes <= 0 || offset < 0 || width <= 0 || offset + width > bytes) {
    codens_set_error("codens_host_bundle_fetch: invalid argument");
    return CODENS_EINVAL;
  }
  const char *d = (const char *)find_mirror((const char *)hosts[0] + 1, (size_t)bytes * count);
  if (!d) {
    codens_set_error("codens_host_bundle_fetch: no such bundle");
    return CODENS_EINVAL;
  }
  for (int i = 0; i < count; i++)
    if (hipMemcpyAsync((char *)hosts[i] + offset, d + (size_t)i * bytes + offset, (size_t)width,
                       hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess) {
      codens_set_error("codens_host_bundle_fetch: device to host copy failed");
      return CODENS_EHIP;
    }
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    codens_set_error("codens_host_bundle_fetch: synchronisation failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}
