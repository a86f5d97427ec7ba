// Shared definitions for the gfx950 HIP implementation of libmat's dense
// GEMM + factor-and-solve path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <unordered_map>

namespace mg {

// ---------------------------------------------------------------- errors --
// Every HIP failure inside the library is turned into a C++ exception at
// the call site; the C-ABI layer converts it into a status code (extension
// entry points) or an assert-like abort (the drop-in libmat symbols, whose
// reference signatures have no error channel, mat.h:328-357).
struct Error {
  int code;
  std::string msg;
};

void set_last_error(const std::string &m);
const char *last_error();

#define MG_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::mg::Error{-1, std::string(#expr) + ": " + hipGetErrorString(e_) \
                                + " (" __FILE__ ":" + std::to_string(__LINE__) + ")"}; \
  } while (0)

#define MG_CHECK(cond, msg)                                                   \
  do {                                                                        \
    if (!(cond)) throw ::mg::Error{-2, std::string(msg)};                     \
  } while (0)

// ------------------------------------------------------------ workspace --
// Device scratch owned by the library, one set per (device, stream):
// named slots that grow on demand and are reused by later calls on the
// same stream (stream order makes reuse safe).  Freed at process exit.
// The slot map is locked, so host threads may look up slots of one stream
// at once; a pointer stays valid until its slot grows, so calls that
// enqueue work on one stream from several threads must still be ordered by
// the caller (the drop-in entry points use a stream per host thread).
struct Workspace {
  std::mutex mu;
  std::unordered_map<std::string, std::pair<void *, size_t>> slots;
  std::unordered_map<std::string, std::pair<void *, size_t>> hslots;  // pinned host
  // fresh (optional) is set when the slot was (re)allocated: its contents
  // are undefined
  void *get(const std::string &name, size_t bytes, bool *fresh = nullptr);
  // Pinned, coherent host memory (hipHostMalloc), for words that kernels
  // of this stream write and the host polls; contents undefined when new.
  void *get_host(const std::string &name, size_t bytes);
  // the slot's buffer, or null when it was never allocated
  void *find(const std::string &name);
  ~Workspace();
};
Workspace &workspace(hipStream_t s);

// One int per stream workspace, zero at creation, set to 1 by any
// multi-workgroup kernel whose bounded wait for another workgroup expired.
// sync_and_check() synchronizes the stream and, if the word is set, clears
// it and throws.
int *sync_error_word(hipStream_t s);
void sync_and_check(hipStream_t s);

// ------------------------------------------------------------ launch timing --
// Opt-in (mat_gpu_timing_enable): a KScope placed around a launch records a
// HIP event pair on the launch stream and files the launch under a kernel
// family together with its algorithmic work (bytes or flops, the family's
// own unit).  Disabled, a KScope costs one branch.
struct KScope {
  KScope(hipStream_t s, const char *family, double work = 0.0);
  ~KScope();
  KScope(const KScope &) = delete;
  KScope &operator=(const KScope &) = delete;
  hipStream_t s;
  long slot;
};
void timing_enable(bool on);
// launches, summed milliseconds and summed work of one family (synchronizes
// on the recorded events)
void timing_query(const char *family, long *launches, double *ms, double *work);

// Library-owned stream of the calling host thread on the current device,
// for the drop-in (synchronous) entry points: threads never share one, so
// their uploads, workspaces and kernels stay apart.  A thread's streams go
// back to a per-device pool when it exits and are handed to later threads.
hipStream_t default_stream();

// Second library-owned stream per device for the look-ahead
// factorizations' trailing updates (LU, Cholesky, QR).  With
// MATGPU_PANEL_CUS=k (default 0: no mask) and a runtime that supports CU
// masks, it leaves k compute units to the caller's stream.
hipStream_t side_stream();
// Wave priority (s_setprio) for kernels launched on stream s: 0 on a side
// stream, 3 elsewhere.  Kernels of the caller's stream carry a
// factorization's critical path (panels, small solves and products); when
// they share SIMDs with the side stream's trailing-update GEMMs, the
// higher priority lets their instructions issue first.
int wave_prio(hipStream_t s);
// Reusable events (timing disabled) of the current device, by index.
hipEvent_t pooled_event(int i);
// Held by a look-ahead factorization (LU, Cholesky, QR) while it enqueues:
// the side stream and the event pool are shared by every factorization on
// a device.
std::mutex &lookahead_mutex();
// Look-ahead block width of the LU (MATGPU_LU_NB, default 256; 0 = none).
int lookahead_block();

// Number of compute units of the current device.
int num_cus();

// ------------------------------------------------------------ strided views
// A matrix operand addressed as p[i*rs + j*cs]; exactly one of rs/cs is 1.
template <typename T>
struct View {
  T *p;
  long rs, cs;
  __host__ __device__ T &operator()(long i, long j) const { return p[i * rs + j * cs]; }
  View sub(long i, long j) const { return View{p + i * rs + j * cs, rs, cs}; }
  View t() const { return View{p, cs, rs}; }
};
template <typename T>
View<T> colmajor(T *p, long ld) { return View<T>{p, 1, ld}; }
template <typename T>
View<T> rowmajor(T *p, long ld) { return View<T>{p, ld, 1}; }
template <typename T>
View<const T> cview(const View<T> &v) { return View<const T>{v.p, v.rs, v.cs}; }

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------ GEMM entry --
// C = alpha * A * B + beta * C on arbitrary strided views (M x K, K x N,
// M x N).  beta == 0 never reads C.  tri = 1 (or true): only elements with
// row >= col of C are written; tri = 2: only row <= col (tiles wholly
// outside the triangle are skipped); tri = 0: all of C.  grid_cap > 0
// limits the launch to that many workgroups, each walking several tiles
// (work meant to run beside a critical path on another stream).
template <typename T>
void gemm(hipStream_t s, long M, long N, long K, T alpha, View<const T> A,
          View<const T> B, T beta, View<T> C, int tri = 0, int grid_cap = 0);

// Below this many multiply-adds (M*N*K) the factorizations run their
// bf16-split updates as native fp32 MFMA GEMMs instead.
constexpr double kLowpMinWork = 2.0e9;

// Low-precision GEMM used by the mixed solvers: C(fp32) += alpha * A * B with
// A, B given in fp32 and multiplied on BF16 MFMA after an exact split into
// `passes` (1, 3 or 6) bf16 terms; accumulation in fp32.  lower_only: only
// row >= col of C is written (tiles above the diagonal are not launched).
// grid_cap > 0 limits the launch to that many workgroups (rounded down to a
// multiple of 8), each walking several tiles: work beside a critical path on
// another stream leaves compute-unit slots for it.
void gemm_bf16x(hipStream_t s, long M, long N, long K, float alpha,
                View<const float> A, View<const float> B, View<float> C,
                int passes, bool lower_only = false, int grid_cap = 0);

// FP8 e4m3 GEMM with one scale per 128x128 tile of each operand (computed
// on the fly), fp32 accumulation: C += alpha * A * B.
void gemm_fp8(hipStream_t s, long M, long N, long K, float alpha,
              View<const float> A, View<const float> B, View<float> C,
              bool lower_only);

// ---------------------------------------------------- precision policy --
// How the trailing updates of a factorization are multiplied.
enum class Mul { Native = 0, Bf16x1 = 1, Bf16x3 = 3, Bf16x6 = 6, Fp8 = 8 };

}  // namespace mg
