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
//    r of its matrix in registers.  A column step finds the pivot with one
//    64-bit key max over the group (DPP / swizzle) and hands the pivot row
//    (LU) or the column (Cholesky) to the group through the wavefront's own
//    LDS (one lane or the column's lanes store, every lane reads float4
//    chunks back); no workgroup barrier is involved.  The Cholesky back
//    substitution stages L through LDS.
//  * Row interchanges are logical: each lane tracks the position of its
//    row; the pivot search picks max |a| with the lowest position on ties
//    (mat__iamax, mat.h:1247-1262); the last column is not pivoted
//    (mat.h:4214) and a pivot below eps skips its elimination step.
//  * Forward substitution is fused into the elimination; back
//    substitution reads U from the owning lanes' registers.
//  * One persistent kernel per (bin, kind) class, so each class gets the
//    registers and occupancy its matrices need: its wavefronts pull tasks
//    (one task = the matrices of one wavefront) from the class's atomic
//    counter.  The launches (a counter reset, then every class kernel
//    depending on it alone, so the classes run side by side) are built
//    once into a hipGraph of explicit kernel nodes and replayed.
//  * Autotuning (coordinate descent, graph replays timed on the actual
//    batch) picks the bins in use (the order thresholds: a small bin can be
//    merged into the next), the wavefronts per workgroup of every class
//    (2, 4 or 8), the resident workgroups per CU and the class order.
#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
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
  unsigned *counter;  // one task counter per (bin, kind) class
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

// Wavefront-level sync for the LDS hand-offs below (one wavefront owns its
// buffers): order the LDS stores before the loads that follow.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_max_key(uint64_t k) {
  const uint32_t lo = __builtin_amdgcn_update_dpp((uint32_t)k, (uint32_t)k, CTRL, 0xf, 0xf, false);
  const uint32_t hi =
      __builtin_amdgcn_update_dpp((uint32_t)(k >> 32), (uint32_t)(k >> 32), CTRL, 0xf, 0xf, false);
  const uint64_t o = ((uint64_t)hi << 32) | lo;
  return o > k ? o : k;
}

// Max of a 64-bit key over each aligned group of NB lanes (every lane of the
// group gets it): DPP within 16-lane rows (xor 1, xor 2, then the half-row
// and row mirrors, which pair every lane with the other half), a swizzle
// for xor 16 and readlanes across the two 32-lane halves.
template <int NB>
__device__ __forceinline__ uint64_t group_max_key(uint64_t k) {
  k = dpp_max_key<0xB1>(k);  // quad_perm(1,0,3,2)
  k = dpp_max_key<0x4E>(k);  // quad_perm(2,3,0,1)
  k = dpp_max_key<0x141>(k);  // row_half_mirror
  if constexpr (NB >= 16) k = dpp_max_key<0x140>(k);  // row_mirror
  if constexpr (NB >= 32) {
    constexpr int XOR16 = 0x1f | (0x10 << 10);  // bit mode: and 0x1f, xor 0x10
    const uint32_t lo = __builtin_amdgcn_ds_swizzle((int)(uint32_t)k, XOR16);
    const uint32_t hi = __builtin_amdgcn_ds_swizzle((int)(uint32_t)(k >> 32), XOR16);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    k = o > k ? o : k;
  }
  if constexpr (NB == 64) {
    const uint64_t a = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(k >> 32), 0) << 32) |
                       __builtin_amdgcn_readlane((uint32_t)k, 0);
    const uint64_t b = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(k >> 32), 32) << 32) |
                       __builtin_amdgcn_readlane((uint32_t)k, 32);
    k = a > b ? a : b;
  }
  return k;
}

// Per-wavefront LDS: two hand-off buffers (by step parity) of 64 row or
// column values plus one value per group, then (Cholesky) the staged L.
constexpr int HB = 80;  // floats per hand-off buffer (64 values + 8 groups + pad)

template <int NB>
__device__ __forceinline__ void store_row(float *dst, const float (&a)[NB], int from) {
  // whole aligned float4 chunks from the one holding `from` (static after
  // unrolling); entries before `from` are written but never read
  bfor<0, NB / 4>([&](auto qc) __attribute__((always_inline)) {
    constexpr int q = decltype(qc)::value;
    if (4 * q + 3 >= from)
      *reinterpret_cast<float4 *>(dst + 4 * q) =
          make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
  });
}

// Partial-pivot LU + solve of the matrices of one task (64/NB per
// wavefront, group g in lanes [g*NB, g*NB+NB), lane r of a group holding row
// r).  The pivot search is one 64-bit key max per column (|a| bits, then the
// lowest position, then the lane); the pivot row reaches the group through
// LDS (one lane stores it, every lane reads it back in float4 chunks).
template <int NB>
__device__ __forceinline__ void group_lu(const BatchP &p, const BatchTask &t, int lane, float *buf) {
  const int gid = lane / NB, r = lane % NB;
  const bool has = gid < t.count;
  const int mi = has ? p.order[t.first + gid] : 0;
  const int n = has ? p.n[mi] : 0;
  const bool active = r < n;
  const float *A = p.A + (has ? p.aoff[mi] : 0);
  float a[NB];
  bfor<0, NB>([&](auto cc) __attribute__((always_inline)) {
    constexpr int c = decltype(cc)::value;
    a[c] = (active && c < n) ? A[r + (long)c * n] : 0.f;
  });
  float bv = active ? p.b[p.voff[mi] + r] : 0.f;
  int pos = r;
  int pl[NB];
  bfor<0, NB>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    // pivot: first position of max |a(:,k)| among positions >= k (a NaN
    // counts as 0); the last column takes the row already at position k
    uint64_t key = 0;
    const bool elig = active && (k + 1 < n ? pos >= k : pos == k);
    if (elig) {
      float v = k + 1 < n ? fabsf(a[k]) : 1.f;
      if (!(v == v)) v = 0.f;
      key = ((uint64_t)__float_as_uint(v) << 32) |
            (uint64_t)((1u << 12) | ((uint32_t)(63 - pos) << 6) | (uint32_t)lane);
    }
    key = group_max_key<NB>(key);
    const int wl = kvalid ? (int)(key & 63) : lane;
    const int wpos = 63 - (int)((key >> 6) & 63);
    pl[k] = wl;
    const bool was_k = (pos == k);
    if (kvalid) {
      if (lane == wl) pos = k;
      else if (was_k) pos = wpos;
    }
    float *hb = buf + (k & 1) * HB;
    float *row = hb + gid * NB;
    if (kvalid && lane == wl) {
      store_row<NB>(row, a, k);
      hb[64 + gid] = bv;
    }
    wave_lds_sync();
    const float yk = hb[64 + gid];
    float u[NB];
    bfor<k / 4, NB / 4>([&](auto qc) __attribute__((always_inline)) {
      constexpr int q = decltype(qc)::value;
      const float4 v = *reinterpret_cast<const float4 *>(row + 4 * q);
      u[4 * q] = v.x;
      u[4 * q + 1] = v.y;
      u[4 * q + 2] = v.z;
      u[4 * q + 3] = v.w;
    });
    const float ukk = u[k];
    const bool upd = kvalid && active && pos > k && !(fabsf(ukk) < p.eps);
    const float l = upd ? a[k] * (1.0f / ukk) : 0.f;
    if (upd) {
      a[k] = l;
      bv = fmaf(-l, yk, bv);
    }
    bfor<k + 1, NB>([&](auto cc) __attribute__((always_inline)) {
      constexpr int c = decltype(cc)::value;
      if (upd) a[c] = fmaf(-l, u[c], a[c]);
    });
  });
  // back substitution: lane at position k owns U(k, :)
  float s = bv;
  bfor_down<0, NB>([&](auto kc) __attribute__((always_inline)) {
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

// Cholesky + solve of the matrices of one task.  Column k: the lanes at or
// below the diagonal store their (unscaled) entries to LDS, every lane reads
// the diagonal and the entries it needs back, scales them by 1/l(k,k)
// itself (the same operation lane c performs on its own entry) and updates
// its row; forward substitution is fused.  L is then staged in LDS for the
// back substitution.
template <int NB>
__device__ __forceinline__ void group_chol(const BatchP &p, const BatchTask &t, int lane, float *buf) {
  const int gid = lane / NB, r = lane % NB;
  const bool has = gid < t.count;
  const int mi = has ? p.order[t.first + gid] : 0;
  const int n = has ? p.n[mi] : 0;
  const bool active = r < n;
  const float *A = p.A + (has ? p.aoff[mi] : 0);
  float a[NB];
  bfor<0, NB>([&](auto cc) __attribute__((always_inline)) {
    constexpr int c = decltype(cc)::value;
    a[c] = (active && c < n && c <= r) ? A[r + (long)c * n] : 0.f;  // lower triangle only
  });
  float bv = active ? p.b[p.voff[mi] + r] : 0.f;
  bool fail = false;
  bfor<0, NB>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    float *hb = buf + (k & 1) * HB;
    float *col = hb + gid * NB;
    if (kvalid && active && r >= k) col[r] = a[k];
    if (kvalid && r == k) hb[64 + gid] = bv;
    wave_lds_sync();
    const float d = col[k];
    if (kvalid && d <= 0.f) fail = true;
    const float lkk = sqrtf(d);
    const float inv = 1.0f / lkk;
    if (kvalid && active) {
      if (r == k) a[k] = lkk;
      else if (r > k) a[k] = a[k] * inv;
    }
    const float lrk = a[k];
    // the rest of the column in float4 chunks, each used as it arrives
    bfor<(k + 1) / 4, NB / 4>([&](auto qc) __attribute__((always_inline)) {
      constexpr int q = decltype(qc)::value;
      const float4 v = *reinterpret_cast<const float4 *>(col + 4 * q);
      const float u[4] = {v.x, v.y, v.z, v.w};
      bfor<0, 4>([&](auto jc) __attribute__((always_inline)) {
        constexpr int c = 4 * q + decltype(jc)::value;
        // unpredicated: in rows r < c this writes entries above the
        // diagonal, and for k >= n or r >= n entries of unused rows and
        // columns; none of them is read again (the staged L is read at and
        // below the diagonal only)
        if constexpr (c > k) a[c] = fmaf(-lrk, u[c - 4 * q] * inv, a[c]);
      });
    });
    // forward substitution L y = b (division by the diagonal, mat.h:4596)
    const float yk = hb[64 + gid] / lkk;
    if (kvalid && active) {
      if (r == k) bv = yk;
      else if (r > k) bv = fmaf(-lrk, yk, bv);
    }
    // one column step at a time: interleaving steps only raises the
    // register count past the file
    __builtin_amdgcn_sched_barrier(0);
  });
  // stage L rows in LDS for L^T x = y
  float *Ls = buf + 2 * HB + gid * NB * NB;
  bfor<0, NB / 4>([&](auto qc) __attribute__((always_inline)) {
    constexpr int q = decltype(qc)::value;
    *reinterpret_cast<float4 *>(Ls + r * NB + 4 * q) =
        make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
  });
  wave_lds_sync();
  // the row registers are free from here on: the diagonal comes from LDS
  const float dg = Ls[r * NB + r];
  float s = bv;
  bfor_down<0, NB>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    const bool kvalid = k < n;
    const float mine = (r == k) ? s / dg : 0.f;
    const float xk = __shfl(mine, gid * NB + k);
    if (kvalid && active && r < k) s = fmaf(-Ls[k * NB + r], xk, s);
    if (kvalid && r == k && !fail) p.x[p.voff[mi] + k] = xk;
  });
  wave_lds_sync();
  if (has && r == 0) p.info[mi] = fail ? -1 : 0;
}

// One kernel per (kind, size bin), so each gets the registers and occupancy
// its matrices need.  Persistent: the wavefronts pull tasks [t0, t1) of
// their class from the class's counter.
constexpr int NCLASS = 8;  // (bin 0..3) x (kind 0..1)
constexpr int class_of(int bin, int kind) { return bin * 2 + kind; }

template <int NB, int KIND>
struct ClassShape {
  // wavefronts per workgroup when not tuned (the autotuner tries 2, 4, 8)
  static constexpr int WAVES = (KIND == 1 && NB == 64) ? 2 : 4;
  // floats per wavefront: hand-off buffers, plus the staged L for Cholesky
  static constexpr int LDSW = 2 * HB + (KIND == 1 ? 64 * NB : 0);
  // wavefronts per SIMD the register allocation must allow (the 64-wide
  // Cholesky would otherwise take one more than half the file)
  static constexpr int MINW = (KIND == 1 && NB == 64) ? 2 : 1;
};

template <int NB, int KIND, int W>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(ClassShape<NB, KIND>::MINW))) void batched_kernel(BatchP p, int t0,
                                                                                  int t1, int cls) {
  using S = ClassShape<NB, KIND>;
  __shared__ __align__(16) float lds[W][S::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  while (true) {
    int task = 0;
    if (lane == 0) task = t0 + (int)atomicAdd(p.counter + cls, 1u);
    task = __shfl(task, 0);
    if (task >= t1) break;
    const BatchTask t = p.tasks[task];
    if constexpr (KIND == 0) group_lu<NB>(p, t, lane, lds[wave]);
    else group_chol<NB>(p, t, lane, lds[wave]);
  }
}

}  // namespace mg

// ------------------------------------------------ no-allocation scratch --
// Device scratch of mat_gpu_batch_solve_noalloc_f32: the batch's orders,
// its matrices in task order, the task queue and the per-class counters,
// for at most NOALLOC_MAX matrices.  Each GPU holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
namespace {
constexpr size_t NOALLOC_MAX = size_t(1) << 18;
constexpr size_t round256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t NA_N = 0;
constexpr size_t NA_ORDER = round256(NA_N + NOALLOC_MAX * sizeof(int));
constexpr size_t NA_TASKS = round256(NA_ORDER + NOALLOC_MAX * sizeof(int));
constexpr size_t NA_COUNTER = round256(NA_TASKS + NOALLOC_MAX * sizeof(mg::BatchTask));
constexpr size_t NA_BYTES = round256(NA_COUNTER + mg::NCLASS * sizeof(unsigned));
}  // namespace
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// ------------------------------------------------------------- plan API --
namespace {

using mg::BatchTask;

struct Plan {
  mg::BatchP p{};
  int *d_order = nullptr;
  BatchTask *d_tasks = nullptr;
  unsigned *d_counter = nullptr;
  int *d_n = nullptr;
  int wg_per_cu = 4;
  // bins in use: bit b = the group of 8 << b lanes (orders up to 8 << b);
  // an order goes to the smallest bin in use that holds it (bin 3 always)
  int bin_mask = 0xF;
  int waves[mg::NCLASS] = {mg::ClassShape<8, 0>::WAVES,  mg::ClassShape<8, 1>::WAVES,
                           mg::ClassShape<16, 0>::WAVES, mg::ClassShape<16, 1>::WAVES,
                           mg::ClassShape<32, 0>::WAVES, mg::ClassShape<32, 1>::WAVES,
                           mg::ClassShape<64, 0>::WAVES, mg::ClassShape<64, 1>::WAVES};
  int cls_t0[mg::NCLASS] = {}, cls_t1[mg::NCLASS] = {};
  int cls_order[mg::NCLASS] = {};  // classes with tasks, in launch order
  int ncls = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuned_ms = 0;
  double untuned_ms = 0;
  int tuned_wg_per_cu = 0;
  int tuned_order = 0;
};

int bin_of(int n, int mask = 0xF) {
  for (int b = 0; b < 3; ++b)
    if (((mask >> b) & 1) && n <= (8 << b)) return b;
  return 3;
}

// Build tasks: bins (NB, kind) -> chunks of 64/NB matrices.  order_mode 0:
// largest bins first; 1: smallest first.  Within a class the matrices keep
// their batch order.  Writes order[count] and tasks[<= count] (a counting
// sort: no allocation) and returns the number of tasks.
int make_tasks_into(size_t count, const int32_t *n, const int8_t *kind, int order_mode,
                    int *order, BatchTask *tasks, Plan &pl) {
  size_t cnt[mg::NCLASS] = {}, off[mg::NCLASS] = {}, pos[mg::NCLASS] = {};
  for (size_t i = 0; i < count; ++i) {
    MG_CHECK(n[i] >= 1 && n[i] <= 64, "batch: matrix order must be in [1, 64]");
    ++cnt[mg::class_of(bin_of(n[i], pl.bin_mask), kind[i] ? 1 : 0)];
  }
  size_t o = 0;
  for (int bi = 0; bi < 4; ++bi) {
    const int b = order_mode == 0 ? 3 - bi : bi;
    for (int k = 0; k < 2; ++k) {
      const int c = mg::class_of(b, k);
      off[c] = pos[c] = o;
      o += cnt[c];
    }
  }
  for (size_t i = 0; i < count; ++i)
    order[pos[mg::class_of(bin_of(n[i], pl.bin_mask), kind[i] ? 1 : 0)]++] = (int)i;
  int nt = 0;
  pl.ncls = 0;
  for (int bi = 0; bi < 4; ++bi) {
    const int b = order_mode == 0 ? 3 - bi : bi;
    const size_t per = 64 / (8 << b);
    for (int k = 0; k < 2; ++k) {
      const int c = mg::class_of(b, k);
      pl.cls_t0[c] = nt;
      if (cnt[c]) pl.cls_order[pl.ncls++] = c;
      for (size_t s = 0; s < cnt[c]; s += per) {
        BatchTask t;
        t.bin = b;
        t.kind = k;
        t.first = (int)(off[c] + s);
        t.count = (int)std::min(per, cnt[c] - s);
        tasks[nt++] = t;
      }
      pl.cls_t1[c] = nt;
    }
  }
  return nt;
}

void make_tasks(size_t count, const int32_t *n, const int8_t *kind, int order_mode,
                std::vector<int> &order, std::vector<BatchTask> &tasks, Plan &pl) {
  order.resize(count);
  tasks.resize(count);
  tasks.resize(make_tasks_into(count, n, kind, order_mode, order.data(), tasks.data(), pl));
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
}

template <int NB, int KIND, int W>
void launch_class_w(hipStream_t s, const Plan &pl, int c) {
  const int t0 = pl.cls_t0[c], t1 = pl.cls_t1[c];
  // resident wavefronts as for wg_per_cu workgroups of four per CU, and no
  // more workgroups than the class has tasks for
  const int grid = std::max(1, std::min((int)mg::ceil_div(t1 - t0, W),
                                        mg::num_cus() * pl.wg_per_cu * 4 / W));
  hipLaunchKernelGGL((mg::batched_kernel<NB, KIND, W>), dim3(grid), dim3(64 * W), 0, s, pl.p, t0,
                     t1, c);
  MG_HIP(hipGetLastError());
}

template <int NB, int KIND>
void launch_class(hipStream_t s, const Plan &pl, int c) {
  switch (pl.waves[c]) {
    case 2: launch_class_w<NB, KIND, 2>(s, pl, c); break;
    case 8: launch_class_w<NB, KIND, 8>(s, pl, c); break;
    default: launch_class_w<NB, KIND, 4>(s, pl, c); break;
  }
}

void launch_one(hipStream_t s, const Plan &pl, int c) {
  switch (c) {
    case mg::class_of(0, 0): launch_class<8, 0>(s, pl, c); break;
    case mg::class_of(0, 1): launch_class<8, 1>(s, pl, c); break;
    case mg::class_of(1, 0): launch_class<16, 0>(s, pl, c); break;
    case mg::class_of(1, 1): launch_class<16, 1>(s, pl, c); break;
    case mg::class_of(2, 0): launch_class<32, 0>(s, pl, c); break;
    case mg::class_of(2, 1): launch_class<32, 1>(s, pl, c); break;
    case mg::class_of(3, 0): launch_class<64, 0>(s, pl, c); break;
    default: launch_class<64, 1>(s, pl, c); break;
  }
}

// The class kernels in launch order, after the counters' reset.
void launch(hipStream_t s, Plan &pl) {
  mg::fill_u32(s, pl.d_counter, mg::NCLASS, 0u);
  for (int i = 0; i < pl.ncls; ++i) launch_one(s, pl, pl.cls_order[i]);
}

// The graph of a plan, built from explicit kernel nodes: the counters'
// reset (a kernel, not a memset node), then every class kernel depending on
// it alone, so the classes run side by side (they share nothing: each pulls
// its own tasks from its own counter and writes its own matrices).  No
// capture, so no event or side stream is involved: the fork/join capture
// this replaces destroyed its events right after instantiation while the
// graph could still refer to them.
__global__ void batch_reset_kernel(unsigned *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

struct ClassLaunch {
  const void *fn;
  int grid, block;
};

template <int NB, int KIND, int W>
ClassLaunch class_launch_w(const Plan &pl, int c) {
  const int t0 = pl.cls_t0[c], t1 = pl.cls_t1[c];
  // the grid of launch_class_w
  const int grid = std::max(1, std::min((int)mg::ceil_div(t1 - t0, W),
                                        mg::num_cus() * pl.wg_per_cu * 4 / W));
  return {reinterpret_cast<const void *>(&mg::batched_kernel<NB, KIND, W>), grid, 64 * W};
}

template <int NB, int KIND>
ClassLaunch class_launch(const Plan &pl, int c) {
  switch (pl.waves[c]) {
    case 2: return class_launch_w<NB, KIND, 2>(pl, c);
    case 8: return class_launch_w<NB, KIND, 8>(pl, c);
    default: return class_launch_w<NB, KIND, 4>(pl, c);
  }
}

ClassLaunch class_launch_of(const Plan &pl, int c) {
  switch (c) {
    case mg::class_of(0, 0): return class_launch<8, 0>(pl, c);
    case mg::class_of(0, 1): return class_launch<8, 1>(pl, c);
    case mg::class_of(1, 0): return class_launch<16, 0>(pl, c);
    case mg::class_of(1, 1): return class_launch<16, 1>(pl, c);
    case mg::class_of(2, 0): return class_launch<32, 0>(pl, c);
    case mg::class_of(2, 1): return class_launch<32, 1>(pl, c);
    case mg::class_of(3, 0): return class_launch<64, 0>(pl, c);
    default: return class_launch<64, 1>(pl, c);
  }
}

void capture(hipStream_t s, Plan &pl) {
  if (pl.exec) MG_HIP(hipGraphExecDestroy(pl.exec));
  if (pl.graph) MG_HIP(hipGraphDestroy(pl.graph));
  pl.exec = nullptr;
  pl.graph = nullptr;
  MG_HIP(hipStreamSynchronize(s));
  MG_HIP(hipGraphCreate(&pl.graph, 0));
  unsigned *ctr = pl.d_counter;
  int ncl = mg::NCLASS;
  void *rargs[] = {&ctr, &ncl};
  hipKernelNodeParams rp{};
  rp.func = reinterpret_cast<void *>(&batch_reset_kernel);
  rp.gridDim = dim3(1);
  rp.blockDim = dim3(64);
  rp.sharedMemBytes = 0;
  rp.kernelParams = rargs;
  rp.extra = nullptr;
  hipGraphNode_t reset;
  MG_HIP(hipGraphAddKernelNode(&reset, pl.graph, nullptr, 0, &rp));
  for (int i = 0; i < pl.ncls; ++i) {
    const int c = pl.cls_order[i];
    const ClassLaunch cl = class_launch_of(pl, c);
    mg::BatchP bp = pl.p;
    int t0 = pl.cls_t0[c], t1 = pl.cls_t1[c], cc = c;
    void *args[] = {&bp, &t0, &t1, &cc};  // copied into the node
    hipKernelNodeParams kp{};
    kp.func = const_cast<void *>(cl.fn);
    kp.gridDim = dim3(cl.grid);
    kp.blockDim = dim3(cl.block);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    hipGraphNode_t node;
    MG_HIP(hipGraphAddKernelNode(&node, pl.graph, &reset, 1, &kp));
  }
  MG_HIP(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
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

// Frees whatever a plan holds (also a half-built one); errors are ignored,
// this runs on error paths too.
void free_plan(Plan *pl) {
  if (!pl) return;
  (void)hipDeviceSynchronize();
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->d_order) (void)hipFree(pl->d_order);
  if (pl->d_tasks) (void)hipFree(pl->d_tasks);
  if (pl->d_counter) (void)hipFree(pl->d_counter);
  if (pl->d_n) (void)hipFree(pl->d_n);
  delete pl;
}
struct PlanFree {
  void operator()(Plan *pl) const { free_plan(pl); }
};

void check_orders(size_t count, const int32_t *n) {
  for (size_t i = 0; i < count; ++i)
    MG_CHECK(n[i] >= 1 && n[i] <= 64, "batch: matrix order must be in [1, 64]");
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
    check_orders(count, n);
    hipStream_t s = (hipStream_t)stream;
    // freed on every error path below, handed to the caller on success
    std::unique_ptr<Plan, PlanFree> pl(new Plan());
    std::vector<int> hn(n, n + count);
    MG_HIP(hipMalloc(&pl->d_n, std::max<size_t>(1, count) * sizeof(int)));
    MG_HIP(hipMemcpyAsync(pl->d_n, hn.data(), count * sizeof(int), hipMemcpyHostToDevice, s));
    MG_HIP(hipMalloc(&pl->d_counter, mg::NCLASS * sizeof(unsigned)));
    pl->p.A = a;
    pl->p.aoff = (const long long *)a_off;
    pl->p.b = b;
    pl->p.x = x;
    pl->p.voff = (const long long *)v_off;
    pl->p.n = pl->d_n;
    pl->p.info = info;
    pl->p.counter = pl->d_counter;
    pl->p.eps = 1e-6f;
    std::vector<int> order;
    std::vector<BatchTask> tasks;
    if (!autotune) {
      make_tasks(count, n, kind, 0, order, tasks, *pl);
      upload_plan(s, *pl, order, tasks);
      pl->wg_per_cu = 4;
      pl->tuned_wg_per_cu = 4;
      capture(s, *pl);
    } else {
      // coordinate descent over the plan's choices, each timed by graph
      // replays on this batch: the bins in use (merging a small bin into
      // the next saves a launch and its tail, at the price of idle lanes),
      // the wavefronts per workgroup of every class, the resident
      // workgroups per CU and the class order; a choice is kept when it is
      // at least 1% faster
      int order_mode = 0;
      auto trial = [&]() {
        make_tasks(count, n, kind, order_mode, order, tasks, *pl);
        upload_plan(s, *pl, order, tasks);
        capture(s, *pl);
        return time_replays(s, *pl, 10);  // 10 replays: 3 let noise pick slower plans
      };
      double best = trial();
      pl->untuned_ms = best;
      auto consider = [&](auto set, auto undo) {
        set();
        const double ms = trial();
        if (ms < 0.99 * best) best = ms;
        else undo();
      };
      for (int mask = 0; mask < 8; ++mask) {
        const int prev = pl->bin_mask, cand = mask | 8;
        if (cand == prev) continue;
        consider([&] { pl->bin_mask = cand; }, [&] { pl->bin_mask = prev; });
      }
      make_tasks(count, n, kind, order_mode, order, tasks, *pl);  // classes in use
      const int ncls = pl->ncls;
      int used[mg::NCLASS];
      for (int i = 0; i < ncls; ++i) used[i] = pl->cls_order[i];
      for (int i = 0; i < ncls; ++i) {
        const int c = used[i];
        for (int wv : {2, 4, 8}) {
          const int prev = pl->waves[c];
          if (wv == prev) continue;
          consider([&] { pl->waves[c] = wv; }, [&] { pl->waves[c] = prev; });
        }
      }
      for (int wg : {1, 2, 8, 4}) {
        const int prev = pl->wg_per_cu;
        if (wg == prev) continue;
        consider([&] { pl->wg_per_cu = wg; }, [&] { pl->wg_per_cu = prev; });
      }
      consider([&] { order_mode = 1; }, [&] { order_mode = 0; });
      make_tasks(count, n, kind, order_mode, order, tasks, *pl);
      upload_plan(s, *pl, order, tasks);
      pl->tuned_wg_per_cu = pl->wg_per_cu;
      pl->tuned_order = order_mode;
      pl->tuned_ms = best;
      capture(s, *pl);
    }
    MG_HIP(hipStreamSynchronize(s));
    *out = reinterpret_cast<mat_gpu_batch_plan *>(pl.release());
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

int mat_gpu_batch_tuning_shape(const mat_gpu_batch_plan *plan, int *bin_mask, int32_t *waves,
                               double *untuned_ms) {
  return guarded([&] {
    MG_CHECK(plan != nullptr, "batch: null plan");
    const Plan *pl = reinterpret_cast<const Plan *>(plan);
    if (bin_mask) *bin_mask = pl->bin_mask;
    if (waves)
      for (int c = 0; c < mg::NCLASS; ++c) waves[c] = pl->waves[c];
    if (untuned_ms) *untuned_ms = pl->untuned_ms;
    return 0;
  });
}

int mat_gpu_batch_destroy(mat_gpu_batch_plan *plan) {
  return guarded([&] {
    if (!plan) return 0;
    free_plan(reinterpret_cast<Plan *>(plan));
    return 0;
  });
}

// The no-allocation entry point's state: the scratch array's device
// address per device (resolved once per device, with it current), and host
// staging of the order and the task queue.  One lock covers all of it and
// is held for a whole call, so calls from several threads take turns.
std::mutex g_na_mu;
constexpr int NA_DEVICES = 64;
unsigned char *g_na_ptr[NA_DEVICES] = {};
int g_na_order[NOALLOC_MAX];
mg::BatchTask g_na_tasks[NOALLOC_MAX];

// caller holds g_na_mu
unsigned char *noalloc_scratch() {
  int dev = 0;
  MG_HIP(hipGetDevice(&dev));
  MG_CHECK(dev >= 0 && dev < NA_DEVICES, "batch noalloc: device index out of range");
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    MG_CHECK(e == hipSuccess && p != nullptr,
             std::string("batch noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = (unsigned char *)p;
  }
  return g_na_ptr[dev];
}

// One launch without a graph (used for per-matrix launch loops too).
int mat_gpu_batch_solve_f32(void *stream, size_t count, const int32_t *n, const int8_t *kind,
                            const float *a, const int64_t *a_off, const float *b, float *x,
                            const int64_t *v_off, int32_t *info) {
  return guarded([&] {
    MG_CHECK(n != nullptr && kind != nullptr, "batch: null argument");
    check_orders(count, n);
    hipStream_t s = (hipStream_t)stream;
    mg::Workspace &ws = mg::workspace(s);
    std::vector<int> order;
    std::vector<BatchTask> tasks;
    Plan pl;
    make_tasks(count, n, kind, 0, order, tasks, pl);
    int *d_n = (int *)ws.get("batch.n", std::max<size_t>(1, count) * 4);
    int *d_order = (int *)ws.get("batch.order", std::max<size_t>(1, order.size()) * 4);
    BatchTask *d_tasks =
        (BatchTask *)ws.get("batch.tasks", std::max<size_t>(1, tasks.size()) * sizeof(BatchTask));
    unsigned *d_counter = (unsigned *)ws.get("batch.counter", mg::NCLASS * sizeof(unsigned));
    std::vector<int> hn(n, n + count);
    MG_HIP(hipMemcpyAsync(d_n, hn.data(), count * 4, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * sizeof(BatchTask),
                          hipMemcpyHostToDevice, s));
    pl.p.A = a;
    pl.p.aoff = (const long long *)a_off;
    pl.p.b = b;
    pl.p.x = x;
    pl.p.voff = (const long long *)v_off;
    pl.p.n = d_n;
    pl.p.info = info;
    pl.p.order = d_order;
    pl.p.tasks = d_tasks;
    pl.p.counter = d_counter;
    pl.p.eps = 1e-6f;
    pl.d_counter = d_counter;
    pl.wg_per_cu = 4;
    launch(s, pl);
    // host vectors must outlive the async copies
    MG_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

// mat_gpu_batch_solve_f32 with its scratch (orders, task queue, counters)
// in the fixed per-GPU array instead of library workspace: nothing is
// allocated, on the device or the host.  Same kernels and launch shape,
// so the results are the same bits.
int mat_gpu_batch_solve_noalloc_f32(void *stream, size_t count, const int32_t *n,
                                    const int8_t *kind, const float *a, const int64_t *a_off,
                                    const float *b, float *x, const int64_t *v_off,
                                    int32_t *info) {
  return guarded([&] {
    MG_CHECK(n != nullptr && kind != nullptr, "batch: null argument");
    MG_CHECK(count <= NOALLOC_MAX,
             "batch noalloc: at most 262144 matrices per call (the size of the fixed scratch "
             "array); use mat_gpu_batch_solve_f32 or a plan for larger batches");
    check_orders(count, n);
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_na_mu);
    unsigned char *base = noalloc_scratch();
    Plan pl;
    const int nt = make_tasks_into(count, n, kind, 0, g_na_order, g_na_tasks, pl);
    int *d_n = (int *)(base + NA_N);
    int *d_order = (int *)(base + NA_ORDER);
    BatchTask *d_tasks = (BatchTask *)(base + NA_TASKS);
    unsigned *d_counter = (unsigned *)(base + NA_COUNTER);
    if (count) {
      MG_HIP(hipMemcpyAsync(d_n, n, count * sizeof(int), hipMemcpyHostToDevice, s));
      MG_HIP(hipMemcpyAsync(d_order, g_na_order, count * sizeof(int), hipMemcpyHostToDevice, s));
      MG_HIP(hipMemcpyAsync(d_tasks, g_na_tasks, (size_t)nt * sizeof(BatchTask),
                            hipMemcpyHostToDevice, s));
    }
    pl.p.A = a;
    pl.p.aoff = (const long long *)a_off;
    pl.p.b = b;
    pl.p.x = x;
    pl.p.voff = (const long long *)v_off;
    pl.p.n = d_n;
    pl.p.info = info;
    pl.p.order = d_order;
    pl.p.tasks = d_tasks;
    pl.p.counter = d_counter;
    pl.p.eps = 1e-6f;
    pl.d_counter = d_counter;
    pl.wg_per_cu = 4;
    launch(s, pl);  // the counters are cleared by a fill kernel first
    // the staging arrays and n are read by the copies until here
    MG_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

}  // extern "C"
