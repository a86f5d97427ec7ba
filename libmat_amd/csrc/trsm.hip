// Left-side triangular solves B := op(T)^-1 B on strided views.
//
// Used for U12 = L11^-1 A12 inside the LU (the reference's TRSM-by-axpy,
// mat.h:4264-4271), for the Cholesky panel A21 L11^-T, and for the
// substitutions of mat_solve / mat_solve_spd / mat_solve_tri* (the scalar
// TRSVs, mat.h:4584-4803).
//
// Recursive blocking: split T at a multiple of 64, solve the leading block,
// update the other half of B with one GEMM (MFMA), solve the trailing
// block.  Leaves (n <= 64) solve one right-hand-side column per thread.
#include <type_traits>

#include "ops.hpp"

namespace mg {

constexpr int TRSM_TB = 64;

constexpr int TRSM_NT = 64;  // one wavefront per workgroup: more workgroups per solve

// One step's update of the remaining rows in the slots [4*C0, 4*(C0+NC)):
// x[s-1] = x[s] - col[s] * xc (the rows move down one slot as they
// update).  The group's LDS reads are all issued before its first update:
// one LDS latency per group, not per four slots.
template <typename T, int C0, int NC>
__device__ __forceinline__ void trsm_group(T (&x)[TRSM_TB], const T *col, T xc) {
  T v[4 * NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if constexpr (sizeof(T) == 4) {
      const float4 q = reinterpret_cast<const float4 *>(col)[C0 + k];
      v[4 * k] = q.x, v[4 * k + 1] = q.y, v[4 * k + 2] = q.z, v[4 * k + 3] = q.w;
    } else {
      const double2 q0 = reinterpret_cast<const double2 *>(col)[2 * (C0 + k)];
      const double2 q1 = reinterpret_cast<const double2 *>(col)[2 * (C0 + k) + 1];
      v[4 * k] = q0.x, v[4 * k + 1] = q0.y, v[4 * k + 2] = q1.x, v[4 * k + 3] = q1.y;
    }
  }
#pragma unroll
  for (int u = 0; u < 4 * NC; ++u) {
    const int sl = 4 * C0 + u;
    if (sl >= 1) x[sl - 1] = fma(-v[u], xc, x[sl]);
  }
}

// Leaves (n <= 64): one thread per right-hand-side column with the column
// in registers and the triangle in LDS in a skewed layout, Sk[c][s] =
// T(c + s, c) (lower) or T(c - s, c) (upper), zero outside the triangle,
// so the step for column c reads one contiguous aligned run.  The column
// loop is a runtime loop: the current element sits in slot 0, each step
// stores it (finished) and moves the remaining rows down one slot while
// updating them, in groups of slots that uniform branches skip once no
// remaining row is in them.  (A fully unrolled 64-step solve is ~9k
// instructions, and executed once per launch it runs from a cold
// instruction cache.)  Upper solves run the same loop from the last row.
// Non-unit leaves divide by the diagonal, as the reference's TRSVs do.
template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(TRSM_NT) void trsm_leaf_kernel(int n, int nrhs, View<const T> Tm,
                                                            View<T> B, int prio) {
  constexpr int TB = TRSM_TB;
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ __align__(16) T Sk[TB * TB];
  {
    // every load issued before the LDS stores (one memory latency, not TB)
    const int sl = threadIdx.x;
    T v[TB];
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      const int r = LOWER ? c + sl : c - sl;
      v[c] = (c < n && r >= 0 && r < n) ? Tm(r, c) : T(0);
    }
#pragma unroll
    for (int c = 0; c < TB; ++c) Sk[c * TB + sl] = v[c];
  }
  __syncthreads();
  const int j = blockIdx.x * TRSM_NT + threadIdx.x;
  if (j >= nrhs) return;
  T x[TB];
#pragma unroll
  for (int r = 0; r < TB; ++r) x[r] = r < n ? B(LOWER ? r : n - 1 - r, j) : T(0);
#pragma unroll 1
  for (int st = 0; st < n; ++st) {
    const int c = LOWER ? st : n - 1 - st;
    const T *col = Sk + c * TB;
    if constexpr (!UNIT) x[0] = x[0] / col[0];
    const T xc = x[0];
    B(c, j) = xc;
    const int live = n - st;  // slots 0 .. live-1 hold this and the remaining rows
    // groups of slots past the remaining rows are skipped (uniform branches)
    if constexpr (sizeof(T) == 4) {
      trsm_group<T, 0, 8>(x, col, xc);
      if (32 < live) trsm_group<T, 8, 8>(x, col, xc);
    } else {
      trsm_group<T, 0, 4>(x, col, xc);
      if (16 < live) trsm_group<T, 4, 4>(x, col, xc);
      if (32 < live) trsm_group<T, 8, 4>(x, col, xc);
      if (48 < live) trsm_group<T, 12, 4>(x, col, xc);
    }
  }
}

// Leaves of 17..128 rows with each right-hand-side column split over G
// lanes of one wavefront, R rows per lane (64/G columns per wavefront, four
// wavefronts per workgroup).  The triangle is staged once
// in LDS in solve order (an upper solve is the lower solve of the reversed
// rows), column-major with a padded stride.  Step c: the lane holding row c
// divides by the diagonal (non-unit), stores the finished element and hands
// it to the other lanes of its column (one cross-lane read); lanes holding
// later rows apply the column's R coefficients of their rows.  Per step a
// lane does R FMAs where the one-lane-per-column leaf does up to 64.
template <typename T, bool LOWER, bool UNIT, int G, int R>
__global__ __launch_bounds__(256) void trsm_split_kernel(int n, int nrhs, View<const T> Tm,
                                                         View<T> B, int prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  constexpr int CW = 64 / G;                   // columns per wavefront
  constexpr int NP = R * G;                    // rows held (solve order)
  constexpr int LS = NP + 16 / (int)sizeof(T);  // padded column stride, 16-byte aligned
  constexpr int VW = 16 / (int)sizeof(T);       // elements per 16-byte LDS read
  // Ls[c * LS + p] = L'(p, c) strictly below the diagonal, zero elsewhere;
  // the diagonal apart in Ds (non-unit solves)
  __shared__ __align__(16) T Ls[NP * LS];
  __shared__ T Ds[NP];
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  auto row = [&](int p) { return LOWER ? p : n - 1 - p; };
  // the index order follows the view's contiguous dimension
  const bool tr_rm = Tm.cs == 1;
  for (int e = t; e < NP * NP; e += 256) {
    const int a = e / NP, b = e % NP;
    const int p = (tr_rm == LOWER) ? a : b, c = (tr_rm == LOWER) ? b : a;
    T v = T(0);
    if (p < n && c < n && p >= c) v = Tm(row(p), row(c));
    if (p == c) {
      Ds[c] = c < n ? v : T(1);  // padding steps divide 0 by 1
      v = T(0);
    }
    Ls[c * LS + p] = v;
  }
  __syncthreads();
  const int jl = l % CW, q = l / CW;
  const int j = (blockIdx.x * 4 + w) * CW + jl;
  if ((blockIdx.x * 4 + w) * CW >= nrhs) return;  // whole wavefront idle
  const bool colok = j < nrhs;
  T x[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int p = R * q + i;
    x[i] = (colok && p < n) ? B(row(p), j) : T(0);
  }
  // this lane's R rows of column c
  auto column = [&](int c, T(&tc)[R]) {
    const T *col = Ls + c * LS + R * q;
#pragma unroll
    for (int k = 0; k < R; k += VW) {
      if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(col + k);
        tc[k] = v.x, tc[k + 1] = v.y, tc[k + 2] = v.z, tc[k + 3] = v.w;
      } else {
        const double2 v = *reinterpret_cast<const double2 *>(col + k);
        tc[k] = v.x, tc[k + 1] = v.y;
      }
    }
  };
  T tc[R], fin[R];
#pragma unroll
  for (int i = 0; i < R; ++i) fin[i] = T(0);
  column(0, tc);
  // runtime loop over the row groups, unrolled over the R steps inside one
  // (static register indices): the loop body stays small enough to run
  // from a warm instruction cache, where R*G unrolled steps would not.
  // No branch per step: steps past n see zero coefficients and broadcast
  // zero, and each finished element is kept in fin[] (one select) and
  // stored at the end.
#pragma unroll 1
  for (int qc = 0; qc < G; ++qc) {
    const bool same = q == qc;
    const int owner = (jl + CW * qc) * 4;  // bpermute byte address of the owner lane
#pragma unroll
    for (int sl = 0; sl < R; ++sl) {
      const int c = R * qc + sl;
      if constexpr (!UNIT) {
        const T d = Ds[c];
        const T xd = x[sl] / d;
        x[sl] = same ? xd : x[sl];
      }
      T xc;
      if constexpr (sizeof(T) == 4) {
        xc = __int_as_float(__builtin_amdgcn_ds_bpermute(owner, __float_as_int(x[sl])));
      } else {
        const long long u = __double_as_longlong(x[sl]);
        const int lo = __builtin_amdgcn_ds_bpermute(owner, (int)(unsigned)u);
        const int hi = __builtin_amdgcn_ds_bpermute(owner, (int)(unsigned)(u >> 32));
        xc = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                              (unsigned)lo));
      }
      // the next column's coefficients while xc is in flight
      T tn[R];
      column(c + 1 < NP ? c + 1 : NP - 1, tn);
      fin[sl] = same ? x[sl] : fin[sl];
      // unconditional: coefficients on or above the diagonal are zero, so
      // rows already solved only take zero updates in registers whose
      // values were kept in fin[]
#pragma unroll
      for (int i = 0; i < R; ++i) x[i] = fma(-tc[i], xc, x[i]);
#pragma unroll
      for (int i = 0; i < R; ++i) tc[i] = tn[i];
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int p = R * q + i;
    if (colok && p < n) B(row(p), j) = fin[i];
  }
}

// C -= A * B with the requested multiply precision.
template <typename T>
static void gemm_sub(hipStream_t s, long M, long N, long K, View<const T> A, View<const T> B,
                     View<T> C, Mul mul) {
  if constexpr (std::is_same<T, float>::value) {
    // small products: native fp32 MFMA is as fast as packing for the split
    // passes (and exact to fp32); the split pays off on large updates only
    if ((mul == Mul::Bf16x1 || mul == Mul::Bf16x3 || mul == Mul::Bf16x6) &&
        (double)M * N * K >= kLowpMinWork) {
      gemm_bf16x(s, M, N, K, -1.0f, A, B, C, (int)mul);
      return;
    }
    if (mul == Mul::Fp8 && (double)M * N * K >= kLowpMinWork) {
      gemm_fp8(s, M, N, K, -1.0f, A, B, C, false);
      return;
    }
  }
  gemm<T>(s, M, N, K, T(-1), A, B, T(1), C, false);
}

template <typename T>
int tri_block_size();
template <typename T>
void tri_invert_blocks(hipStream_t s, bool lower, bool unit, int n, View<const T> Tm, T *D);

// Blocked recursion whose leaves apply precomputed inverses of the TB x TB
// diagonal blocks with one MFMA GEMM each.  In-place is safe: a leaf's
// GEMM has at most one tile along the triangle's dimension, so every
// workgroup reads exactly the block of B it overwrites, and reads it all
// before its epilogue writes.
template <typename T>
static void trsm_inv_rec(hipStream_t s, bool lower, int n, int nrhs, View<const T> Tm, View<T> B,
                         const T *D, int b0, int tb, Mul mul) {
  if (n <= tb) {
    View<const T> Dv{D + (long)b0 * tb * tb, tb, 1};
    gemm<T>(s, n, nrhs, n, T(1), Dv, cview(B), T(0), B, false);
    return;
  }
  int h = ((n / 2 + tb - 1) / tb) * tb;
  if (h >= n) h = n - tb;
  const int r = n - h;
  if (lower) {
    trsm_inv_rec<T>(s, true, h, nrhs, Tm, B, D, b0, tb, mul);
    gemm_sub<T>(s, r, nrhs, h, Tm.sub(h, 0), cview(B), B.sub(h, 0), mul);
    trsm_inv_rec<T>(s, true, r, nrhs, Tm.sub(h, h), B.sub(h, 0), D, b0 + h / tb, tb, mul);
  } else {
    trsm_inv_rec<T>(s, false, r, nrhs, Tm.sub(h, h), B.sub(h, 0), D, b0 + h / tb, tb, mul);
    gemm_sub<T>(s, h, nrhs, r, Tm.sub(0, h), cview(B.sub(h, 0)), B, mul);
    trsm_inv_rec<T>(s, false, h, nrhs, Tm, B, D, b0, tb, mul);
  }
}

template <typename T>
void trsm_left(hipStream_t s, bool lower, bool unit, int n, int nrhs, View<const T> Tm, View<T> B,
               Mul mul) {
  if (n <= 0 || nrhs <= 0) return;
  // narrow right-hand sides (the factorizations' critical-path solves): two
  // split leaves and one GEMM, fewer and shorter launches than inverting
  // the diagonal blocks first; wide ones keep the inverse-block GEMMs,
  // which run on MFMA
  if (n > 128 && n <= 256 && nrhs <= 256) {
    const int h = 128, r = n - h;
    if (lower) {
      trsm_left<T>(s, true, unit, h, nrhs, Tm, B, mul);
      gemm_sub<T>(s, r, nrhs, h, Tm.sub(h, 0), cview(B), B.sub(h, 0), mul);
      trsm_left<T>(s, true, unit, r, nrhs, Tm.sub(h, h), B.sub(h, 0), mul);
    } else {
      trsm_left<T>(s, false, unit, r, nrhs, Tm.sub(h, h), B.sub(h, 0), mul);
      gemm_sub<T>(s, h, nrhs, r, Tm.sub(0, h), cview(B.sub(h, 0)), B, mul);
      trsm_left<T>(s, false, unit, h, nrhs, Tm, B, mul);
    }
    return;
  }
  const int prio = wave_prio(s);
  if (n > 16 && (n <= 64 ? nrhs <= 4096 : n <= 128 && nrhs <= 256)) {
    KScope ks(s, "trsm");
    const dim3 blk(256);
#define MG_SPLIT(G_, R_)                                                                            \
  {                                                                                                 \
    const dim3 grid(ceil_div(nrhs, 4 * (64 / G_)));                                               \
    if (lower && unit) hipLaunchKernelGGL((trsm_split_kernel<T, true, true, G_, R_>), grid, blk, 0, s, n, nrhs, Tm, B, prio);   \
    else if (lower) hipLaunchKernelGGL((trsm_split_kernel<T, true, false, G_, R_>), grid, blk, 0, s, n, nrhs, Tm, B, prio);     \
    else if (unit) hipLaunchKernelGGL((trsm_split_kernel<T, false, true, G_, R_>), grid, blk, 0, s, n, nrhs, Tm, B, prio);      \
    else hipLaunchKernelGGL((trsm_split_kernel<T, false, false, G_, R_>), grid, blk, 0, s, n, nrhs, Tm, B, prio);              \
  }
    if (n <= 32) MG_SPLIT(4, 8) else if (n <= 64) MG_SPLIT(8, 8) else MG_SPLIT(16, 8)
#undef MG_SPLIT
    MG_HIP(hipGetLastError());
    return;
  }
  if (n <= TRSM_TB && nrhs <= 4096) {
    // small: one launch of the register-resident leaf
    dim3 grid(ceil_div(nrhs, TRSM_NT)), blk(TRSM_NT);
    KScope ks(s, "trsm");
    if (lower && unit) hipLaunchKernelGGL((trsm_leaf_kernel<T, true, true>), grid, blk, 0, s, n, nrhs, Tm, B, prio);
    else if (lower) hipLaunchKernelGGL((trsm_leaf_kernel<T, true, false>), grid, blk, 0, s, n, nrhs, Tm, B, prio);
    else if (unit) hipLaunchKernelGGL((trsm_leaf_kernel<T, false, true>), grid, blk, 0, s, n, nrhs, Tm, B, prio);
    else hipLaunchKernelGGL((trsm_leaf_kernel<T, false, false>), grid, blk, 0, s, n, nrhs, Tm, B, prio);
    MG_HIP(hipGetLastError());
    return;
  }
  const int tb = tri_block_size<T>();
  const size_t dsz = (size_t)((n + tb - 1) / tb) * tb * tb * sizeof(T);
  T *D = (T *)workspace(s).get("trsm.dinv", dsz);
  tri_invert_blocks<T>(s, lower, unit, n, Tm, D);
  trsm_inv_rec<T>(s, lower, n, nrhs, Tm, B, D, 0, tb, mul);
}

template void trsm_left<float>(hipStream_t, bool, bool, int, int, View<const float>, View<float>, Mul);
template void trsm_left<double>(hipStream_t, bool, bool, int, int, View<const double>, View<double>, Mul);

}  // namespace mg
