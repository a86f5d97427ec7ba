// Batched factor-and-solve of many small matrices of uneven size (3..64),
// one right-hand side each: general matrices by partial-pivot LU
// (mat_plu + mat_solve semantics, mat.h:4201-4288, 4807-4837) and SPD
// matrices by Cholesky (mat_chol + mat_solve_spd, mat.h:6031-6060,
// 4839-4871; -1 per matrix that is not positive definite, x untouched).
//
// MI355X design:
//  * Size bins: a matrix of order n is handled by a group of NB lanes,
//    NB = max(8, next power of two >= n) in {8, 16, 32, 64}, so a
//    wavefront factors 64/NB matrices at once.  Lane r of a group holds row
//    r of its matrix in registers; column steps exchange values with
//    in-group shuffles, so a factorization never touches LDS or a barrier
//    (the Cholesky back substitution stages L through LDS).
//  * Row interchanges are logical: each lane tracks the position of its
//    row; the pivot search picks max |a| with the lowest position on ties
//    (mat__iamax, mat.h:1247-1262); the last column is not pivoted
//    (mat.h:4214) and a pivot below eps skips its elimination step.
//  * Forward substitution is fused into the elimination; back
//    substitution reads U from the owning lanes' registers.
//  * One persistent launch: wavefronts pull tasks (one task = the matrices
//    of one wavefront, same bin and kind) from an atomic counter, largest
//    bins first.  The launch (counter reset + kernel) is captured once in a
//    hipGraph and replayed.
//  * Autotuning picks the number of resident workgroups per CU and the
//    task order by timing graph replays on the actual batch.
#include <algorithm>
#include <chrono>
#include <vector>

#include "mat_gpu.h"
#include "ops.hpp"

namespace mg {

struct BatchTask {
  int bin;    // 0..3 -> NB = 8 << bin
  int kind;   // 0 general, 1 SPD
  int first;  // into order[]
  int count;  // matrices in this task (<= 64/NB)
};

struct BatchP {
  const float *A;
  const long long *aoff;
  const float *b;
  float *x;
  const long long *voff;
  const int *n;
  int *info;
  const int *order;
  const BatchTask *tasks;
  int ntasks;
  unsigned *counter;
  float eps;
};

template <int B, int E, typename F>
__device__ __forceinline__ void bfor(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    bfor<B + 1, E>(f);
  }
}
template <int B, int E, typename F>
__device__ __forceinline__ void bfor_down(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, E - 1>{});
    bfor_down<B, E - 1>(f);
  }
}

template <int NB>
__device__ void group_lu(const BatchP &p, const BatchTask &t, int lane) {
  const int gid = lane / NB, r = lane % NB, gbase = gid * NB;
  const bool has = gid < t.count;
  const int mi = has ? p.order[t.first + gid] : 0;
  const int n = has ? p.n[mi] : 0;
  const bool active = r < n;
  const float *A = p.A + (has ? p.aoff[mi] : 0);
  float a[NB];
  bfor<0, NB>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    a[c] = (active && c < n) ? A[r + (long)c * n] : 0.f;
  });
  float bv = active ? p.b[p.voff[mi] + r] : 0.f;
  int pos = r;
  int pl[NB];
  bfor<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    // pivot: first position of max |a(:,k)| among positions >= k; the last
    // column takes the row already at position k
    float v;
    if (k + 1 < n) v = (active && pos >= k) ? fabsf(a[k]) : -1.f;
    else v = (active && pos == k) ? 1.f : -1.f;
    int vp = pos, vl = lane;
#pragma unroll
    for (int off = NB / 2; off >= 1; off >>= 1) {
      const float ov = __shfl_xor(v, off);
      const int op = __shfl_xor(vp, off), ol = __shfl_xor(vl, off);
      if (ov > v || (ov == v && op < vp)) {
        v = ov;
        vp = op;
        vl = ol;
      }
    }
    const int wl = kvalid ? vl : lane;
    const int wpos = vp;
    pl[k] = wl;
    const bool was_k = (pos == k);
    if (kvalid) {
      if (lane == wl) pos = k;
      else if (was_k) pos = wpos;
    }
    const float ukk = __shfl(a[k], wl);
    const float yk = __shfl(bv, wl);
    const bool upd = kvalid && active && pos > k && !(fabsf(ukk) < p.eps);
    const float l = upd ? a[k] * (1.0f / ukk) : 0.f;
    if (upd) {
      a[k] = l;
      bv = fmaf(-l, yk, bv);
    }
#pragma unroll
    for (int c = k + 1; c < NB; ++c) {
      const float uc = __shfl(a[c], wl);
      if (upd) a[c] = fmaf(-l, uc, a[c]);
    }
  });
  // back substitution: lane at position k owns U(k, :)
  float s = bv;
  bfor_down<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    const int wl = pl[k];
    const float mine = (lane == wl && kvalid) ? s / a[k] : 0.f;
    const float xk = __shfl(mine, wl);
    if (kvalid && active && pos < k) s = fmaf(-a[k], xk, s);
    if (kvalid && lane == wl) p.x[p.voff[mi] + k] = xk;
  });
  if (has && r == 0) p.info[mi] = 0;
}

template <int NB>
__device__ void group_chol(const BatchP &p, const BatchTask &t, int lane, float *lds) {
  const int gid = lane / NB, r = lane % NB, gbase = gid * NB;
  const bool has = gid < t.count;
  const int mi = has ? p.order[t.first + gid] : 0;
  const int n = has ? p.n[mi] : 0;
  const bool active = r < n;
  const float *A = p.A + (has ? p.aoff[mi] : 0);
  float a[NB];
  bfor<0, NB>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    a[c] = (active && c < n && c <= r) ? A[r + (long)c * n] : 0.f;  // lower triangle only
  });
  float bv = active ? p.b[p.voff[mi] + r] : 0.f;
  bool fail = false;
  bfor<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    const float d = __shfl(a[k], gbase + k);
    if (kvalid && d <= 0.f) fail = true;
    const float lkk = sqrtf(d);
    const float inv = 1.0f / lkk;
    if (kvalid && active) {
      if (r == k) a[k] = lkk;
      else if (r > k) a[k] = a[k] * inv;
    }
    const float lrk = a[k];
#pragma unroll
    for (int c = k + 1; c < NB; ++c) {
      const float lck = __shfl(a[k], gbase + c);
      if (kvalid && active && r >= c) a[c] = fmaf(-lrk, lck, a[c]);
    }
    // forward substitution L y = b (division by the diagonal, mat.h:4596)
    const float bk = __shfl(bv, gbase + k);
    const float yk = bk / lkk;
    if (kvalid && active) {
      if (r == k) bv = yk;
      else if (r > k) bv = fmaf(-lrk, yk, bv);
    }
  });
  // stage L rows in LDS for L^T x = y
  float *Ls = lds + gid * NB * NB;
  bfor<0, NB>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    Ls[r * NB + c] = a[c];
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float s = bv;
  bfor_down<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    const float mine = (r == k) ? s / a[k] : 0.f;
    const float xk = __shfl(mine, gbase + k);
    if (kvalid && active && r < k) s = fmaf(-Ls[k * NB + r], xk, s);
    if (kvalid && r == k && !fail) p.x[p.voff[mi] + k] = xk;
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (has && r == 0) p.info[mi] = fail ? -1 : 0;
}

constexpr int BATCH_WAVES = 4;

__global__ __launch_bounds__(64 * BATCH_WAVES) void batched_kernel(BatchP p) {
  __shared__ float lds[BATCH_WAVES][64 * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  while (true) {
    int task = 0;
    if (lane == 0) task = (int)atomicAdd(p.counter, 1u);
    task = __shfl(task, 0);
    if (task >= p.ntasks) break;
    const BatchTask t = p.tasks[task];
    if (t.kind == 0) {
      switch (t.bin) {
        case 0: group_lu<8>(p, t, lane); break;
        case 1: group_lu<16>(p, t, lane); break;
        case 2: group_lu<32>(p, t, lane); break;
        default: group_lu<64>(p, t, lane); break;
      }
    } else {
      switch (t.bin) {
        case 0: group_chol<8>(p, t, lane, lds[wave]); break;
        case 1: group_chol<16>(p, t, lane, lds[wave]); break;
        case 2: group_chol<32>(p, t, lane, lds[wave]); break;
        default: group_chol<64>(p, t, lane, lds[wave]); break;
      }
    }
  }
}

}  // namespace mg

// ------------------------------------------------------------- plan API --
namespace {

using mg::BatchTask;

struct Plan {
  mg::BatchP p{};
  int *d_order = nullptr;
  BatchTask *d_tasks = nullptr;
  unsigned *d_counter = nullptr;
  int *d_n = nullptr;
  int grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuned_ms = 0;
  int tuned_wg_per_cu = 0;
  int tuned_order = 0;
};

int bin_of(int n) { return n <= 8 ? 0 : (n <= 16 ? 1 : (n <= 32 ? 2 : 3)); }

// Build tasks: bins (NB, kind) -> chunks of 64/NB matrices.  order_mode 0:
// largest bins first; 1: smallest first.
void make_tasks(size_t count, const int32_t *n, const int8_t *kind, int order_mode,
                std::vector<int> &order, std::vector<BatchTask> &tasks) {
  std::vector<int> idx[4][2];
  for (size_t i = 0; i < count; ++i) {
    MG_CHECK(n[i] >= 1 && n[i] <= 64, "batch: matrix order must be in [1, 64]");
    idx[bin_of(n[i])][kind[i] ? 1 : 0].push_back((int)i);
  }
  order.clear();
  tasks.clear();
  for (int bi = 0; bi < 4; ++bi) {
    const int b = order_mode == 0 ? 3 - bi : bi;
    const int per = 64 / (8 << b);
    for (int k = 0; k < 2; ++k) {
      auto &v = idx[b][k];
      for (size_t s = 0; s < v.size(); s += per) {
        BatchTask t;
        t.bin = b;
        t.kind = k;
        t.first = (int)order.size();
        t.count = (int)std::min<size_t>(per, v.size() - s);
        for (int q = 0; q < t.count; ++q) order.push_back(v[s + q]);
        tasks.push_back(t);
      }
    }
  }
}

void upload_plan(hipStream_t s, Plan &pl, const std::vector<int> &order,
                 const std::vector<BatchTask> &tasks) {
  if (pl.d_order) MG_HIP(hipFree(pl.d_order));
  if (pl.d_tasks) MG_HIP(hipFree(pl.d_tasks));
  MG_HIP(hipMalloc(&pl.d_order, std::max<size_t>(1, order.size()) * sizeof(int)));
  MG_HIP(hipMalloc(&pl.d_tasks, std::max<size_t>(1, tasks.size()) * sizeof(BatchTask)));
  MG_HIP(hipMemcpyAsync(pl.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s));
  MG_HIP(hipMemcpyAsync(pl.d_tasks, tasks.data(), tasks.size() * sizeof(BatchTask),
                        hipMemcpyHostToDevice, s));
  pl.p.order = pl.d_order;
  pl.p.tasks = pl.d_tasks;
  pl.p.ntasks = (int)tasks.size();
}

void launch(hipStream_t s, Plan &pl) {
  mg::fill_u32(s, pl.d_counter, 1, 0u);
  hipLaunchKernelGGL(mg::batched_kernel, dim3(pl.grid), dim3(64 * mg::BATCH_WAVES), 0, s, pl.p);
  MG_HIP(hipGetLastError());
}

void capture(hipStream_t s, Plan &pl) {
  if (pl.exec) MG_HIP(hipGraphExecDestroy(pl.exec));
  if (pl.graph) MG_HIP(hipGraphDestroy(pl.graph));
  hipStream_t cs;
  MG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  MG_HIP(hipStreamSynchronize(s));
  MG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  launch(cs, pl);
  MG_HIP(hipStreamEndCapture(cs, &pl.graph));
  MG_HIP(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
  MG_HIP(hipStreamDestroy(cs));
}

double time_replays(hipStream_t s, Plan &pl, int reps) {
  MG_HIP(hipGraphLaunch(pl.exec, s));
  MG_HIP(hipStreamSynchronize(s));
  hipEvent_t e0, e1;
  MG_HIP(hipEventCreate(&e0));
  MG_HIP(hipEventCreate(&e1));
  MG_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) MG_HIP(hipGraphLaunch(pl.exec, s));
  MG_HIP(hipEventRecord(e1, s));
  MG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MG_HIP(hipEventElapsedTime(&ms, e0, e1));
  MG_HIP(hipEventDestroy(e0));
  MG_HIP(hipEventDestroy(e1));
  return ms / reps;
}

template <typename F>
int guarded(F &&f) {
  try {
    mg::set_last_error("");
    return f();
  } catch (const mg::Error &e) {
    mg::set_last_error(e.msg);
    return e.code < 0 ? e.code - 100 : -100;
  }
}

}  // namespace

extern "C" {

int mat_gpu_batch_create_f32(void *stream, size_t count, const int32_t *n, const int8_t *kind,
                             const float *a, const int64_t *a_off, const float *b, float *x,
                             const int64_t *v_off, int32_t *info, int autotune,
                             mat_gpu_batch_plan **out) {
  return guarded([&] {
    MG_CHECK(out != nullptr && n != nullptr && kind != nullptr, "batch: null argument");
    hipStream_t s = (hipStream_t)stream;
    Plan *pl = new Plan();
    std::vector<int> hn(n, n + count);
    MG_HIP(hipMalloc(&pl->d_n, std::max<size_t>(1, count) * sizeof(int)));
    MG_HIP(hipMemcpyAsync(pl->d_n, hn.data(), count * sizeof(int), hipMemcpyHostToDevice, s));
    MG_HIP(hipMalloc(&pl->d_counter, sizeof(unsigned)));
    pl->p.A = a;
    pl->p.aoff = (const long long *)a_off;
    pl->p.b = b;
    pl->p.x = x;
    pl->p.voff = (const long long *)v_off;
    pl->p.n = pl->d_n;
    pl->p.info = info;
    pl->p.counter = pl->d_counter;
    pl->p.eps = 1e-6f;
    const int cus = mg::num_cus();
    std::vector<int> order;
    std::vector<BatchTask> tasks;
    if (!autotune) {
      make_tasks(count, n, kind, 0, order, tasks);
      upload_plan(s, *pl, order, tasks);
      pl->grid = cus * 4;
      pl->tuned_wg_per_cu = 4;
      capture(s, *pl);
    } else {
      double best = 1e30;
      int best_wg = 4, best_order = 0;
      for (int om = 0; om < 2; ++om) {
        make_tasks(count, n, kind, om, order, tasks);
        upload_plan(s, *pl, order, tasks);
        for (int wg : {1, 2, 4, 8}) {
          pl->grid = cus * wg;
          capture(s, *pl);
          const double ms = time_replays(s, *pl, 3);
          if (ms < best) {
            best = ms;
            best_wg = wg;
            best_order = om;
          }
        }
      }
      make_tasks(count, n, kind, best_order, order, tasks);
      upload_plan(s, *pl, order, tasks);
      pl->grid = cus * best_wg;
      pl->tuned_wg_per_cu = best_wg;
      pl->tuned_order = best_order;
      pl->tuned_ms = best;
      capture(s, *pl);
    }
    MG_HIP(hipStreamSynchronize(s));
    *out = reinterpret_cast<mat_gpu_batch_plan *>(pl);
    return 0;
  });
}

int mat_gpu_batch_run(void *stream, mat_gpu_batch_plan *plan) {
  return guarded([&] {
    MG_CHECK(plan != nullptr, "batch: null plan");
    Plan *pl = reinterpret_cast<Plan *>(plan);
    hipStream_t s = (hipStream_t)stream;
    MG_HIP(hipGraphLaunch(pl->exec, s));
    return 0;
  });
}

int mat_gpu_batch_tuning(const mat_gpu_batch_plan *plan, int *wg_per_cu, int *order_mode,
                         double *tuned_ms) {
  return guarded([&] {
    MG_CHECK(plan != nullptr, "batch: null plan");
    const Plan *pl = reinterpret_cast<const Plan *>(plan);
    if (wg_per_cu) *wg_per_cu = pl->tuned_wg_per_cu;
    if (order_mode) *order_mode = pl->tuned_order;
    if (tuned_ms) *tuned_ms = pl->tuned_ms;
    return 0;
  });
}

int mat_gpu_batch_destroy(mat_gpu_batch_plan *plan) {
  return guarded([&] {
    if (!plan) return 0;
    Plan *pl = reinterpret_cast<Plan *>(plan);
    MG_HIP(hipDeviceSynchronize());
    if (pl->exec) MG_HIP(hipGraphExecDestroy(pl->exec));
    if (pl->graph) MG_HIP(hipGraphDestroy(pl->graph));
    if (pl->d_order) MG_HIP(hipFree(pl->d_order));
    if (pl->d_tasks) MG_HIP(hipFree(pl->d_tasks));
    if (pl->d_counter) MG_HIP(hipFree(pl->d_counter));
    if (pl->d_n) MG_HIP(hipFree(pl->d_n));
    delete pl;
    return 0;
  });
}

// One launch without a graph (used for per-matrix launch loops too).
int mat_gpu_batch_solve_f32(void *stream, size_t count, const int32_t *n, const int8_t *kind,
                            const float *a, const int64_t *a_off, const float *b, float *x,
                            const int64_t *v_off, int32_t *info) {
  return guarded([&] {
    hipStream_t s = (hipStream_t)stream;
    mg::Workspace &ws = mg::workspace(s);
    std::vector<int> order;
    std::vector<BatchTask> tasks;
    make_tasks(count, n, kind, 0, order, tasks);
    int *d_n = (int *)ws.get("batch.n", std::max<size_t>(1, count) * 4);
    int *d_order = (int *)ws.get("batch.order", std::max<size_t>(1, order.size()) * 4);
    BatchTask *d_tasks =
        (BatchTask *)ws.get("batch.tasks", std::max<size_t>(1, tasks.size()) * sizeof(BatchTask));
    unsigned *d_counter = (unsigned *)ws.get("batch.counter", 4);
    std::vector<int> hn(n, n + count);
    MG_HIP(hipMemcpyAsync(d_n, hn.data(), count * 4, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * sizeof(BatchTask),
                          hipMemcpyHostToDevice, s));
    Plan pl;
    pl.p.A = a;
    pl.p.aoff = (const long long *)a_off;
    pl.p.b = b;
    pl.p.x = x;
    pl.p.voff = (const long long *)v_off;
    pl.p.n = d_n;
    pl.p.info = info;
    pl.p.order = d_order;
    pl.p.tasks = d_tasks;
    pl.p.ntasks = (int)tasks.size();
    pl.p.counter = d_counter;
    pl.p.eps = 1e-6f;
    pl.d_counter = d_counter;
    pl.grid = std::max(1, std::min((int)tasks.size(), mg::num_cus() * 4));
    launch(s, pl);
    // host vectors must outlive the async copies
    MG_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

}  // extern "C"
