// Host runtime: error strings, per-stream device workspaces, the library's
// own stream for the synchronous drop-in entry points.
#include <array>
#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace mg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

void *Workspace::get(const std::string &name, size_t bytes, bool *fresh) {
  if (bytes == 0) bytes = 1;
  if (fresh) *fresh = false;
  std::lock_guard<std::mutex> lk(mu);
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

void *Workspace::get_host(const std::string &name, size_t bytes) {
  if (bytes == 0) bytes = 1;
  std::lock_guard<std::mutex> lk(mu);
  auto it = hslots.find(name);
  if (it != hslots.end() && it->second.second >= bytes) return it->second.first;
  if (it != hslots.end()) {
    // growing: queued work on this stream may still write the old buffer
    MG_HIP(hipDeviceSynchronize());
    MG_HIP(hipHostFree(it->second.first));
    hslots.erase(it);
  }
  void *p = nullptr;
  const size_t rounded = (bytes + 255) / 256 * 256;
  MG_HIP(hipHostMalloc(&p, rounded, hipHostMallocCoherent));
  hslots[name] = {p, rounded};
  return p;
}

void *Workspace::find(const std::string &name) {
  std::lock_guard<std::mutex> lk(mu);
  auto it = slots.find(name);
  return it == slots.end() ? nullptr : it->second.first;
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
  std::lock_guard<std::mutex> lk(ws.mu);
  auto it = ws.slots.find("sync.err");
  if (it != ws.slots.end()) return (int *)it->second.first;
  void *p = nullptr;
  MG_HIP(hipMalloc(&p, 256));
  // once, at creation (never inside a capture), before any kernel can see it
  MG_HIP(hipMemset(p, 0, sizeof(int)));
  MG_HIP(hipStreamSynchronize(nullptr));
  ws.slots["sync.err"] = {p, 256};
  return (int *)p;
}

void sync_and_check(hipStream_t s) {
  int *word = (int *)workspace(s).find("sync.err");
  if (!word) {
    MG_HIP(hipStreamSynchronize(s));
    return;
  }
  int h = 0;
  MG_HIP(hipMemcpyAsync(&h, word, sizeof(int), hipMemcpyDeviceToHost, s));
  MG_HIP(hipStreamSynchronize(s));
  if (h != 0) {
    MG_HIP(hipMemsetAsync(word, 0, sizeof(int), s));
    MG_HIP(hipStreamSynchronize(s));
    throw Error{-5, "a multi-workgroup kernel timed out waiting for another workgroup"};
  }
}

namespace {
struct TimeRec {
  std::string family;
  hipEvent_t e0, e1;
  double work;
};
std::mutex g_tm_mutex;
std::atomic<bool> g_tm_on{false};
std::vector<TimeRec> g_tm_recs;
std::vector<hipEvent_t> g_tm_pool;

hipEvent_t tm_event() {
  if (!g_tm_pool.empty()) {
    hipEvent_t e = g_tm_pool.back();
    g_tm_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  MG_HIP(hipEventCreate(&e));
  return e;
}
}  // namespace

KScope::KScope(hipStream_t s_, const char *family, double work) : s(s_), slot(-1) {
  if (!g_tm_on.load(std::memory_order_relaxed)) return;
  std::lock_guard<std::mutex> lk(g_tm_mutex);
  TimeRec r{family, tm_event(), tm_event(), work};
  MG_HIP(hipEventRecord(r.e0, s));
  slot = (long)g_tm_recs.size();
  g_tm_recs.push_back(r);
}

KScope::~KScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_tm_mutex);
  if (slot < (long)g_tm_recs.size()) (void)hipEventRecord(g_tm_recs[slot].e1, s);
}

void timing_enable(bool on) {
  std::lock_guard<std::mutex> lk(g_tm_mutex);
  if (on) {
    MG_HIP(hipDeviceSynchronize());  // recorded events may still be pending
    for (auto &r : g_tm_recs) {
      g_tm_pool.push_back(r.e0);
      g_tm_pool.push_back(r.e1);
    }
    g_tm_recs.clear();
  }
  g_tm_on.store(on);
}

void timing_query(const char *family, long *launches, double *ms, double *work) {
  std::lock_guard<std::mutex> lk(g_tm_mutex);
  long c = 0;
  double t = 0.0, w = 0.0;
  for (auto &r : g_tm_recs) {
    if (r.family != family) continue;
    MG_HIP(hipEventSynchronize(r.e1));
    float e = 0.0f;
    MG_HIP(hipEventElapsedTime(&e, r.e0, r.e1));
    ++c;
    t += e;
    w += r.work;
  }
  if (launches) *launches = c;
  if (ms) *ms = t;
  if (work) *work = w;
}

// every device's side stream, for wave_prio (no device query per launch)
static std::array<std::atomic<hipStream_t>, 64> &side_streams() {
  static std::array<std::atomic<hipStream_t>, 64> a{};
  return a;
}

hipStream_t side_stream() {
  static std::mutex m;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(m);
  auto it = streams.find(dev);
  if (it != streams.end()) return it->second;
  hipStream_t st = nullptr;
  // measured on MI355X: masking CUs off the trailing-update stream slows
  // the fp32 factorization (its GEMMs lose the CUs) and gains nothing for
  // the others, so the mask is opt-in
  int keep = 0;
  if (const char *e = getenv("MATGPU_PANEL_CUS")) keep = atoi(e);
  const int cus = num_cus();
  if (keep > 0 && keep < cus) {
    // mask off the first `keep` CUs of the side stream
    std::vector<uint32_t> mask((cus + 31) / 32, 0u);
    for (int c = keep; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
    if (hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
      (void)hipGetLastError();
      st = nullptr;
    }
  }
  if (!st) MG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  streams[dev] = st;
  for (auto &slot : side_streams())
    if (slot.load() == nullptr) {
      slot.store(st);
      break;
    }
  return st;
}

int wave_prio(hipStream_t s) {
  for (auto &slot : side_streams())
    if (slot.load() == s && s != nullptr) return 0;
  return 3;
}

hipEvent_t pooled_event(int i) {
  static std::mutex m;
  static std::map<int, std::vector<hipEvent_t>> pools;
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(m);
  auto &v = pools[dev];
  while ((int)v.size() <= i) {
    hipEvent_t e;
    MG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    v.push_back(e);
  }
  return v[i];
}

std::mutex &lookahead_mutex() {
  static std::mutex m;
  return m;
}

int lookahead_block() {
  static int nb = [] {
    const char *e = getenv("MATGPU_LU_NB");
    return e ? atoi(e) : 256;
  }();
  return nb;
}

namespace {
// Streams of exited host threads, per device, for later threads to reuse
// (with their workspaces).  Leaked at process exit, like the workspaces.
std::mutex g_pool_mu;
std::map<int, std::vector<hipStream_t>> *g_pool = new std::map<int, std::vector<hipStream_t>>();

struct ThreadStreams {
  std::map<int, hipStream_t> by_dev;
  ~ThreadStreams() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (auto &kv : by_dev) (*g_pool)[kv.first].push_back(kv.second);
  }
};
thread_local ThreadStreams t_streams;
}  // namespace

hipStream_t default_stream() {
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  auto it = t_streams.by_dev.find(dev);
  if (it != t_streams.by_dev.end()) return it->second;
  hipStream_t s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto &free_list = (*g_pool)[dev];
    if (!free_list.empty()) {
      s = free_list.back();
      free_list.pop_back();
    }
  }
  if (!s) MG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  t_streams.by_dev[dev] = s;
  return s;
}

int num_cus() {
  int dev = 0, n = 0;
  MG_HIP(hipGetDevice(&dev));
  MG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
  return n;
}

}  // namespace mg
