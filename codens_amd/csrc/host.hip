// Host-array mirrors for callers that hold their fields in host memory (the
// Fortran drop-in routines): one device buffer per host array (or per bundle
// of host arrays), found by the host address, grown when the caller's array
// is larger than before.  Only the window a call names is copied -- one
// block's slice of each variable -- so a multi-block caller moves each block
// once per call, not the whole array.
//
// Residency (codens_host_residency(1); off by default).  Off, every window a
// call names moves in before its kernels and back after them.  On, the
// library records, per host array, the windows whose device copy is current
// and, of those, the ones whose device copy is newer than the host's:
//  - a window moving in whose device copy is current is not copied;
//  - a window moving back is not copied: it becomes device-newer;
//  - codens_host_sync(a) copies a's device-newer windows back; the host
//    calls it before its own code reads a;
//  - codens_host_modified(a) forgets a's device copies; the host calls it
//    after its own code writes a.  It fails while a has device-newer windows
//    (the host wrote over values it never read).
// A window moving in first brings back every device-newer window of the
// same host array that it shares bytes with, so what it sends is
// current; a window becoming device-newer in one device buffer first brings
// back, then forgets, the overlapping windows another buffer holds of that
// array (a bundle slot and a plain mirror of one array).  A stage whose host
// code reads and writes none of its arrays then moves nothing over PCIe.
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace {
// Window of `rows` runs of `width` bytes, `pitch` apart, from byte `offset`.
struct Win {
  size_t offset, width, pitch, rows;
  size_t end() const { return offset + (rows - 1) * pitch + width; }
  bool same(const Win &o) const {
    return offset == o.offset && width == o.width && rows == o.rows && (rows == 1 || pitch == o.pitch);
  }
  // shared bytes: run by run (a block's window of a multi-variable array is
  // a few runs, one per variable); by span when there are very many runs
  bool overlaps(const Win &o) const {
    if (!(offset < o.end() && o.offset < end())) return false;
    if (rows * o.rows > 4096) return true;
    for (size_t i = 0; i < rows; i++) {
      size_t a0 = offset + i * pitch, a1 = a0 + width;
      for (size_t j = 0; j < o.rows; j++) {
        size_t b0 = o.offset + j * o.pitch;
        if (a0 < b0 + o.width && b0 < a1) return true;
      }
    }
    return false;
  }
};

bool make_win(int64_t offset, int64_t width, int64_t pitch, int32_t rows, Win &w) {
  if (offset < 0 || width <= 0 || rows <= 0 || (rows > 1 && pitch < width)) return false;
  w = Win{(size_t)offset, (size_t)width, (size_t)(rows > 1 ? pitch : width), (size_t)rows};
  return true;
}

using Key = std::pair<int, const void *>;  // (device, host address); a bundle: first array's address + 1

struct Mirror {
  void *dev = nullptr;
  size_t bytes = 0;
  std::vector<const void *> members;  // a bundle's host arrays, slot s at s * slot bytes
  size_t slot = 0;
};

// A window of host array `host` whose copy in mirror `key` (slot `slot`) is
// current; newer: newer than the host's.
struct Track {
  Key key;
  int slot;
  Win w;
  bool newer;
};

std::mutex g_mu;
std::map<Key, Mirror> g_mirrors;
std::multimap<Key, Track> g_tracks;  // keyed by (device, host array)
bool g_resident = false;
unsigned long long g_h2d = 0, g_d2h = 0;

// addresses of a window's first byte in the host array and in the mirror
char *dev_at(const Mirror &m, int slot, const Win &w) { return (char *)m.dev + (size_t)slot * m.slot + w.offset; }

hipError_t copy_win(char *dst, const char *src, const Win &w, hipMemcpyKind kind, hipStream_t st) {
  (kind == hipMemcpyHostToDevice ? g_h2d : g_d2h) += (unsigned long long)w.width * w.rows;
  if (w.rows == 1) return hipMemcpyAsync(dst, src, w.width, kind, st);
  return hipMemcpy2DAsync(dst, w.pitch, src, w.pitch, w.width, w.rows, kind, st);
}

hipError_t move_out(const void *host, const Track &t, hipStream_t st) {
  auto it = g_mirrors.find(t.key);
  if (it == g_mirrors.end()) return hipErrorInvalidValue;
  return copy_win((char *)host + t.w.offset, dev_at(it->second, t.slot, t.w), t.w, hipMemcpyDeviceToHost, st);
}

// device-newer windows of host array `host` back to the host
hipError_t sync_array(int dev, const void *host, hipStream_t st) {
  auto r = g_tracks.equal_range({dev, host});
  for (auto it = r.first; it != r.second; ++it)
    if (it->second.newer) {
      hipError_t e = move_out(host, it->second, st);
      if (e != hipSuccess) return e;
      it->second.newer = false;
    }
  return hipSuccess;
}

// window w of host array `host`, slot `slot` of mirror `key`, host -> device
hipError_t track_in(const Key &key, int slot, const void *host, const Win &w, hipStream_t st) {
  const Mirror &m = g_mirrors[key];
  if (!g_resident) return copy_win(dev_at(m, slot, w), (const char *)host + w.offset, w, hipMemcpyHostToDevice, st);
  auto r = g_tracks.equal_range({key.first, host});
  for (auto it = r.first; it != r.second; ++it)
    if (it->second.key == key && it->second.slot == slot && it->second.w.same(w)) return hipSuccess;
  for (auto it = r.first; it != r.second; ++it)
    if (it->second.newer && it->second.w.overlaps(w)) {
      hipError_t e = move_out(host, it->second, st);
      if (e != hipSuccess) return e;
      it->second.newer = false;
    }
  hipError_t e = copy_win(dev_at(m, slot, w), (const char *)host + w.offset, w, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) g_tracks.insert({{key.first, host}, Track{key, slot, w, false}});
  return e;
}

// window w of host array `host` (slot `slot` of mirror `key`) after a kernel
// wrote it: device -> host, or, resident, marked device-newer
hipError_t track_out(const Key &key, int slot, void *host, const Win &w, hipStream_t st) {
  const Mirror &m = g_mirrors[key];
  if (!g_resident) return copy_win((char *)host + w.offset, dev_at(m, slot, w), w, hipMemcpyDeviceToHost, st);
  auto r = g_tracks.equal_range({key.first, host});
  bool found = false;
  for (auto it = r.first; it != r.second;) {
    Track &t = it->second;
    bool mine = t.key == key && t.slot == slot;
    if (mine && t.w.same(w)) {
      t.newer = found = true;
    } else if (t.w.overlaps(w)) {
      if (mine) {
        t.newer = true;  // its device copy is current and now newer on the overlap
      } else {           // another buffer's copy: its other bytes back, then stale
        if (t.newer) {
          hipError_t e = move_out(host, t, st);
          if (e != hipSuccess) return e;
        }
        it = g_tracks.erase(it);
        continue;
      }
    }
    ++it;
  }
  if (!found) g_tracks.insert({{key.first, host}, Track{key, slot, w, true}});
  return hipSuccess;
}

// Forget every window held in mirror `key`, bringing the device-newer ones
// back first (the buffer is about to be replaced).
hipError_t drop_mirror_tracks(const Key &key, hipStream_t st) {
  for (auto it = g_tracks.begin(); it != g_tracks.end();) {
    if (it->second.key == key) {
      if (it->second.newer) {
        hipError_t e = move_out(it->first.second, it->second, st);
        if (e != hipSuccess) return e;
      }
      it = g_tracks.erase(it);
    } else {
      ++it;
    }
  }
  return hipStreamSynchronize(st);
}

// the mirror of `key`, at least `bytes` long (caller holds g_mu)
void *mirror_of(const Key &key, size_t bytes, hipStream_t st, bool *fresh = nullptr) {
  Mirror &m = g_mirrors[key];
  if (fresh) *fresh = m.bytes < bytes;
  if (m.bytes < bytes) {
    if (m.dev) {
      if (drop_mirror_tracks(key, st) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return nullptr;
      (void)hipFree(m.dev);
      m.dev = nullptr;
      m.bytes = 0;
    }
    if (hipMalloc(&m.dev, bytes) != hipSuccess) return nullptr;
    m.bytes = bytes;
  }
  return m.dev;
}

int current_device() {
  int dev = 0;
  return hipGetDevice(&dev) == hipSuccess ? dev : -1;
}
}  // namespace

extern "C" int codens_host_mirror(const void *host, int64_t bytes, int64_t offset, int64_t width,
                                  int64_t pitch, int32_t rows, int32_t to_device, void **dev_out,
                                  void *stream) {
  Win w;
  if (!host || bytes <= 0 || !dev_out || !make_win(offset, width, pitch, rows, w) ||
      w.end() > (size_t)bytes || to_device < 0 || to_device > 2) {
    codens_set_error("codens_host_mirror: invalid argument or window outside the array");
    return CODENS_EINVAL;
  }
  int dev = current_device();
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  bool fresh = false;
  Key key{dev, host};
  char *d = dev < 0 ? nullptr : (char *)mirror_of(key, (size_t)bytes, st, &fresh);
  if (!d) {
    codens_set_error("codens_host_mirror: device allocation of %lld bytes failed", (long long)bytes);
    return CODENS_EHIP;
  }
  hipError_t e = hipSuccess;
  if (to_device == 1) {
    e = track_in(key, 0, host, w, st);
  } else if (to_device == 2 && fresh) {  // constant array: whole array, once
    g_h2d += (unsigned long long)bytes;
    e = hipMemcpyAsync(d, host, (size_t)bytes, hipMemcpyHostToDevice, st);
  }
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
  int dev = current_device();
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mirrors.find({dev, host});
  if (it == g_mirrors.end() || it->second.bytes < w.end() || !it->second.members.empty()) {
    codens_set_error("codens_host_fetch: no device mirror of this host array covering the window");
    return CODENS_EINVAL;
  }
  if (track_out({dev, host}, 0, host, w, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    codens_set_error("codens_host_fetch: device to host copy failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

// Packs `count` host arrays of `bytes` each into consecutive slots of one
// device bundle (the tq and fd bundles of the C-ABI, slot stride = the
// arrays' length); the window [offset, offset+width) of each is copied.  A
// bundle is found by its first array; one named with other arrays than
// before gives its windows back first and starts anew.
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
  int dev = current_device();
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  Key key{dev, (const char *)hosts[0] + 1};
  {
    Mirror &m = g_mirrors[key];
    std::vector<const void *> mem(hosts, hosts + count);
    if (m.members != mem || m.slot != (size_t)bytes) {
      if (drop_mirror_tracks(key, st) != hipSuccess) {
        codens_set_error("codens_host_bundle: device to host copy failed");
        return CODENS_EHIP;
      }
      m.members = mem;
      m.slot = (size_t)bytes;
    }
  }
  char *d = dev < 0 ? nullptr : (char *)mirror_of(key, (size_t)bytes * count, st);
  if (!d) {
    codens_set_error("codens_host_bundle: device allocation failed");
    return CODENS_EHIP;
  }
  Win w{(size_t)offset, (size_t)width, (size_t)width, 1};
  if (to_device)
    for (int i = 0; i < count; i++)
      if (track_in(key, i, hosts[i], w, st) != hipSuccess) {
        codens_set_error("codens_host_bundle: host to device copy failed");
        return CODENS_EHIP;
      }
  *dev_out = d;
  return CODENS_OK;
}

// The window of every slot of a bundle made by codens_host_bundle back to
// its host array (or, resident, marked device-newer), then synchronises.
extern "C" int codens_host_bundle_fetch(const void *const *hosts, int32_t count, int64_t bytes, int64_t offset,
                                        int64_t width, void *stream) {
  if (!hosts || count <= 0 || bytes <= 0 || offset < 0 || width <= 0 || offset + width > bytes) {
    codens_set_error("codens_host_bundle_fetch: invalid argument");
    return CODENS_EINVAL;
  }
  int dev = current_device();
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  Key key{dev, (const char *)hosts[0] + 1};
  auto it = g_mirrors.find(key);
  if (it == g_mirrors.end() || it->second.members != std::vector<const void *>(hosts, hosts + count) ||
      it->second.slot != (size_t)bytes) {
    codens_set_error("codens_host_bundle_fetch: no such bundle");
    return CODENS_EINVAL;
  }
  Win w{(size_t)offset, (size_t)width, (size_t)width, 1};
  for (int i = 0; i < count; i++)
    if (track_out(key, i, (void *)hosts[i], w, st) != hipSuccess) {
      codens_set_error("codens_host_bundle_fetch: device to host copy failed");
      return CODENS_EHIP;
    }
  if (hipStreamSynchronize(st) != hipSuccess) {
    codens_set_error("codens_host_bundle_fetch: synchronisation failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

extern "C" int codens_host_residency(int32_t on, int32_t *previous, void *stream) {
  if (on != 0 && on != 1) {
    codens_set_error("codens_host_residency: 0 (copy every window both ways) or 1 (resident), not %d", on);
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  if (previous) *previous = g_resident ? 1 : 0;
  if (g_resident && !on) {  // every device-newer window back; nothing tracked
    for (auto &kv : g_tracks)
      if (kv.second.newer && move_out(kv.first.second, kv.second, st) != hipSuccess) {
        codens_set_error("codens_host_residency: device to host copy failed");
        return CODENS_EHIP;
      }
    g_tracks.clear();
    if (hipStreamSynchronize(st) != hipSuccess) {
      codens_set_error("codens_host_residency: synchronisation failed");
      return CODENS_EHIP;
    }
  }
  g_resident = on == 1;
  return CODENS_OK;
}

extern "C" int codens_host_sync(const void *host, void *stream) {
  if (!host) {
    codens_set_error("codens_host_sync: null array");
    return CODENS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  if (sync_array(current_device(), host, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    codens_set_error("codens_host_sync: device to host copy failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

extern "C" int codens_host_sync_all(void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int dev = current_device();
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_tracks)
    if (kv.first.first == dev && kv.second.newer) {
      if (move_out(kv.first.second, kv.second, st) != hipSuccess) {
        codens_set_error("codens_host_sync_all: device to host copy failed");
        return CODENS_EHIP;
      }
      kv.second.newer = false;
    }
  if (hipStreamSynchronize(st) != hipSuccess) {
    codens_set_error("codens_host_sync_all: synchronisation failed");
    return CODENS_EHIP;
  }
  return CODENS_OK;
}

extern "C" int codens_host_modified(const void *host) {
  if (!host) {
    codens_set_error("codens_host_modified: null array");
    return CODENS_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto r = g_tracks.equal_range({current_device(), host});
  for (auto it = r.first; it != r.second; ++it)
    if (it->second.newer) {
      codens_set_error("codens_host_modified: the array has device-newer windows the host never read "
                       "(codens_host_sync first)");
      return CODENS_EINVAL;
    }
  g_tracks.erase(r.first, r.second);
  return CODENS_OK;
}

extern "C" int codens_host_traffic(int64_t *h2d, int64_t *d2h) {
  if (!h2d || !d2h) {
    codens_set_error("codens_host_traffic: null argument");
    return CODENS_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  *h2d = (int64_t)g_h2d;
  *d2h = (int64_t)g_d2h;
  return CODENS_OK;
}
