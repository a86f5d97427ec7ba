// Host runtime: error strings, per-stream device workspaces, the library's
// own stream for the synchronous drop-in entry points.
#include <map>

#include "common.hpp"

namespace mg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

void *Workspace::get(const std::string &name, size_t bytes) {
  if (bytes == 0) bytes = 1;
  auto it = slots.find(name);
  if (it != slots.end() && it->second.second >= bytes) return it->second.first;
  if (it != slots.end()) {
    // growing: the old buffer may still be used by queued work on this stream
    MG_HIP(hipDeviceSynchronize());
    MG_HIP(hipFree(it->second.first));
    slots.erase(it);
  }
  void *p = nullptr;
  const size_t rounded = (bytes + 255) / 256 * 256;
  MG_HIP(hipMalloc(&p, rounded));
  slots[name] = {p, rounded};
  return p;
}

Workspace::~Workspace() {
  // process teardown: the runtime may already be gone; leave memory to the driver
}

static std::mutex g_ws_mutex;
static std::map<std::pair<int, hipStream_t>, Workspace *> g_ws;

Workspace &workspace(hipStream_t s) {
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mutex);
  auto key = std::make_pair(dev, s);
  auto it = g_ws.find(key);
  if (it != g_ws.end()) return *it->second;
  Workspace *w = new Workspace();
  g_ws[key] = w;
  return *w;
}

hipStream_t default_stream() {
  static std::mutex m;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(m);
  auto it = streams.find(dev);
  if (it != streams.end()) return it->second;
  hipStream_t s;
  MG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  streams[dev] = s;
  return s;
}

int num_cus() {
  int dev = 0, n = 0;
  MG_HIP(hipGetDevice(&dev));
  MG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
  return n;
}

}  // namespace mg
