// Host runtime: error strings, per-stream device workspaces, the library's
// own stream for the synchronous drop-in entry points.
#include <map>

#include "common.hpp"

namespace mg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

void *Workspace::get(const std::string &name, size_t bytes, bool *fresh) {
  if (bytes == 0) bytes = 1;
  if (fresh) *fresh = false;
  auto it = slots.find(name);
  if (it != slots.end() && it->second.second >= bytes) return it->second.first;
  if (fresh) *fresh = true;
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

int *sync_error_word(hipStream_t s) {
  Workspace &ws = workspace(s);
  auto it = ws.slots.find("sync.err");
  if (it != ws.slots.end()) return (int *)it->second.first;
  int *p = (int *)ws.get("sync.err", sizeof(int));
  MG_HIP(hipMemset(p, 0, sizeof(int)));  // once, at creation (never inside a capture)
  return p;
}

void sync_and_check(hipStream_t s) {
  Workspace &ws = workspace(s);
  auto it = ws.slots.find("sync.err");
  if (it == ws.slots.end()) {
    MG_HIP(hipStreamSynchronize(s));
    return;
  }
  int h = 0;
  MG_HIP(hipMemcpyAsync(&h, it->second.first, sizeof(int), hipMemcpyDeviceToHost, s));
  MG_HIP(hipStreamSynchronize(s));
  if (h != 0) {
    MG_HIP(hipMemsetAsync(it->second.first, 0, sizeof(int), s));
    MG_HIP(hipStreamSynchronize(s));
    throw Error{-5, "a multi-workgroup kernel timed out waiting for another workgroup"};
  }
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
