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

// ---------------------------------------- block solves with inverses ----
// X := L^-1 B for a unit lower triangle L of nt <= 256 rows whose 32 x 32
// diagonal blocks come with their inverses (the tournament leaves write
// them, linv + 1024 i for block i), both L and B row-major.  Block by block,
// in one workgroup per 16 right-hand-side columns:
//   Y_i = B_i - L(i, 0 : 32 i) X(0 : 32 i),   X_i = L_ii^-1 Y_i,
// so the chain is nt / 32 block steps with workgroup barriers instead of nt
// dependent row steps.  The product of each step is split over the four
// wavefronts by K (partials summed in wavefront order, deterministic); a
// lane owns 2 rows x 4 columns of it.  Block row i + 1 of L is loaded into
// registers while step i computes and goes to LDS once step i's product is
// done; L_ii^-1 likewise one step ahead.  LDS stays under 64 KB, so a
// workgroup fits on a compute unit beside two split-BF16 GEMM workgroups of
// the side stream.
constexpr int LV_R = 256;           // rows of the largest triangle
constexpr int LV_C = 16;            // right-hand-side columns per workgroup
constexpr int LV_XLD = LV_C + 4;    // LDS row stride of X and the partials
constexpr int LV_LBC = LV_R - 32;   // widest block row of L (224 columns)
constexpr int LV_LBLD = LV_LBC + 4;
constexpr int LV_LSLD = 36;

template <bool VEC>
__device__ __forceinline__ float4 lv_load4(const float *p, int valid) {
  if (VEC && valid >= 4) return *reinterpret_cast<const float4 *>(p);
  float4 v = {0.0f, 0.0f, 0.0f, 0.0f};
  if (valid > 0) v.x = p[0];
  if (valid > 1) v.y = p[1];
  if (valid > 2) v.z = p[2];
  if (valid > 3) v.w = p[3];
  return v;
}

template <bool VEC>
__global__ __launch_bounds__(256) void trsm_linv_kernel(int nt, int nrhs, const float *L, long ldl,
                                                        const float *Li, float *B, long ldb,
                                                        int prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ __align__(16) float Xs[LV_R * LV_XLD];
  __shared__ __align__(16) float Ps[3][32 * LV_XLD];  // partials of wavefronts 1..3, then Y
  __shared__ __align__(16) float LB[32 * LV_LBLD];  // block row i of L
  __shared__ __align__(16) float Ls[32 * LV_LSLD];  // L_ii^-1
  const int t = threadIdx.x, wv = t >> 6, l = t & 63, rg = l >> 2, cg = l & 3;
  const int c0 = blockIdx.x * LV_C;
  const int nbk = (nt + 31) / 32;
  // B rows [0, nt) (zero past nt and past nrhs)
  for (int e = t; e < nbk * 32 * (LV_C / 4); e += 256) {
    const int r = e / (LV_C / 4), c4 = (e % (LV_C / 4)) * 4;
    const float4 v = r < nt ? lv_load4<VEC>(B + (long)r * ldb + c0 + c4, nrhs - c0 - c4)
                            : float4{0.0f, 0.0f, 0.0f, 0.0f};
    *reinterpret_cast<float4 *>(&Xs[r * LV_XLD + c4]) = v;
  }
  // prefetch registers: a block row of L (32 rows x up to 224 columns,
  // float4 e = t + 256 u is row e / 56, columns 4 (e % 56) ..) and one
  // float4 of an inverse block (row t / 8, columns 4 (t % 8) ..)
  constexpr int NU = 32 * LV_LBC / 4 / 256;  // 7
  float4 pl[NU], pi;
  auto fetch_l = [&](int ib) {
    const int cols = 32 * ib;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = t + 256 * u, r = e / (LV_LBC / 4), c4 = (e % (LV_LBC / 4)) * 4;
      pl[u] = (c4 < cols && 32 * ib + r < nt) ? lv_load4<VEC>(L + (long)(32 * ib + r) * ldl + c4, 4)
                                              : float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
  };
  auto commit_l = [&](int ib) {
    const int cols = 32 * ib;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = t + 256 * u, r = e / (LV_LBC / 4), c4 = (e % (LV_LBC / 4)) * 4;
      if (c4 < cols) *reinterpret_cast<float4 *>(&LB[r * LV_LBLD + c4]) = pl[u];
    }
  };
  auto fetch_i = [&](int ib) { pi = *reinterpret_cast<const float4 *>(Li + (long)ib * 1024 + 4 * t); };
  auto commit_i = [&]() { *reinterpret_cast<float4 *>(&Ls[(t >> 3) * LV_LSLD + 4 * (t & 7)]) = pi; };
  fetch_i(0);
  commit_i();
  if (nbk > 1) fetch_l(1);
  // this lane's rows / columns in the Y and X phases
  const int yr = t >> 3, yc = (t & 7) * 2;
  for (int i = 0; i < nbk; ++i) {
    // Xs of the earlier blocks (and B) and block row i in place; the
    // previous step's readers of Ls are past this barrier
    __syncthreads();
    if (i > 0) commit_i();  // L_ii^-1, fetched during the previous step
    if (i + 1 < nbk) fetch_i(i + 1);
    // partial products over this wavefront's quarter of k in [0, 32 i)
    const int kq = 8 * i;  // = 32 i / 4
    const int kb = wv * kq;
    float acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][c] = 0.0f;
    for (int k = kb; k < kb + kq; k += 4) {
      const float4 l0 = *reinterpret_cast<const float4 *>(&LB[(2 * rg) * LV_LBLD + k]);
      const float4 l1 = *reinterpret_cast<const float4 *>(&LB[(2 * rg + 1) * LV_LBLD + k]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const float4 xv = *reinterpret_cast<const float4 *>(&Xs[(k + kk) * LV_XLD + 4 * cg]);
        const float a0 = kk == 0 ? l0.x : kk == 1 ? l0.y : kk == 2 ? l0.z : l0.w;
        const float a1 = kk == 0 ? l1.x : kk == 1 ? l1.y : kk == 2 ? l1.z : l1.w;
        acc[0][0] = fmaf(a0, xv.x, acc[0][0]);
        acc[0][1] = fmaf(a0, xv.y, acc[0][1]);
        acc[0][2] = fmaf(a0, xv.z, acc[0][2]);
        acc[0][3] = fmaf(a0, xv.w, acc[0][3]);
        acc[1][0] = fmaf(a1, xv.x, acc[1][0]);
        acc[1][1] = fmaf(a1, xv.y, acc[1][1]);
        acc[1][2] = fmaf(a1, xv.z, acc[1][2]);
        acc[1][3] = fmaf(a1, xv.w, acc[1][3]);
      }
    }
    if (wv > 0) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
        *reinterpret_cast<float4 *>(&Ps[wv - 1][(2 * rg + a) * LV_XLD + 4 * cg]) =
            float4{acc[a][0], acc[a][1], acc[a][2], acc[a][3]};
    }
    __syncthreads();
    // every reader of LB is done: block row i + 1 goes in, i + 2 is loaded
    if (i + 1 < nbk) {
      commit_l(i + 1);
      if (i + 2 < nbk) fetch_l(i + 2);
    }
    // Y_i = B_i - (partials in wavefront order), by wavefront 0 from its own
    // partial in registers, into Ps[0]
    if (wv == 0) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int row = 2 * rg + a;
        const float4 b = *reinterpret_cast<const float4 *>(&Xs[(32 * i + row) * LV_XLD + 4 * cg]);
        float4 sm = float4{acc[a][0], acc[a][1], acc[a][2], acc[a][3]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float4 pq = *reinterpret_cast<const float4 *>(&Ps[q][row * LV_XLD + 4 * cg]);
          sm.x += pq.x;
          sm.y += pq.y;
          sm.z += pq.z;
          sm.w += pq.w;
        }
        *reinterpret_cast<float4 *>(&Ps[0][row * LV_XLD + 4 * cg]) =
            float4{b.x - sm.x, b.y - sm.y, b.z - sm.z, b.w - sm.w};
      }
    }
    __syncthreads();
    // X_i = L_ii^-1 Y_i
    {
      float o0 = 0.0f, o1 = 0.0f;
#pragma unroll 8
      for (int j = 0; j < 32; ++j) {
        const float lv = Ls[yr * LV_LSLD + j];
        const float2 y = *reinterpret_cast<const float2 *>(&Ps[0][j * LV_XLD + yc]);
        o0 = fmaf(lv, y.x, o0);
        o1 = fmaf(lv, y.y, o1);
      }
      *reinterpret_cast<float2 *>(&Xs[(32 * i + yr) * LV_XLD + yc]) = float2{o0, o1};
    }
  }
  __syncthreads();
  for (int e = t; e < nbk * 32 * (LV_C / 4); e += 256) {
    const int r = e / (LV_C / 4), c4 = (e % (LV_C / 4)) * 4;
    if (r >= nt) continue;
    const float4 v = *reinterpret_cast<const float4 *>(&Xs[r * LV_XLD + c4]);
    float *dst = B + (long)r * ldb + c0 + c4;
    const int valid = nrhs - c0 - c4;
    if (VEC && valid >= 4) {
      *reinterpret_cast<float4 *>(dst) = v;
    } else {
      if (valid > 0) dst[0] = v.x;
      if (valid > 1) dst[1] = v.y;
      if (valid > 2) dst[2] = v.z;
      if (valid > 3) dst[3] = v.w;
    }
  }
}

void trsm_lower_unit_linv(hipStream_t s, int nt, int nrhs, View<const float> L, const float *linv,
                          View<float> B) {
  if (nt <= 0 || nrhs <= 0) return;
  MG_CHECK(nt <= LV_R, "trsm_lower_unit_linv: at most 256 rows");
  MG_CHECK(L.cs == 1 && B.cs == 1, "trsm_lower_unit_linv: row-major L and B");
  MG_CHECK((uintptr_t)linv % 16 == 0, "trsm_lower_unit_linv: inverses must be 16-byte aligned");
  const bool vec = (uintptr_t)L.p % 16 == 0 && L.rs % 4 == 0 && (uintptr_t)B.p % 16 == 0 &&
                   B.rs % 4 == 0;
  KScope ks(s, "trsm");
  const dim3 grid(ceil_div(nrhs, LV_C)), blk(256);
  if (vec)
    hipLaunchKernelGGL(trsm_linv_kernel<true>, grid, blk, 0, s, nt, nrhs, L.p, L.rs, linv, B.p, B.rs,
                       wave_prio(s));
  else
    hipLaunchKernelGGL(trsm_linv_kernel<false>, grid, blk, 0, s, nt, nrhs, L.p, L.rs, linv, B.p,
                       B.rs, wave_prio(s));
  MG_HIP(hipGetLastError());
}

template void trsm_left<float>(hipStream_t, bool, bool, int, int, View<const float>, View<float>, Mul);
template void trsm_left<double>(hipStream_t, bool, bool, int, int, View<const double>, View<double>, Mul);

}  // namespace mg
