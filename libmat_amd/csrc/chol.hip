// Lower Cholesky factorization A = L L^T (mat_chol, mat.h:6062-6095).
//
// Reference semantics (x86 path mat__chol_scalar_generic_, mat.h:6031-6060):
// only the lower triangle of A is read; column j's pivot is
// d = a(j,j) - sum_k l(j,k)^2, the factorization fails (-1) at the first
// d <= 0, l(j,j) = sqrt(d) and the column below is multiplied by 1/l(j,j).
//
// MI355X design: the full-precision driver (potrf_window) walks 512-column
// windows of 64-column blocks.  Per block: one GEMM brings the block's
// columns up to date with the window's earlier panels, then one launch
// factors the 64 x 64 diagonal block and solves the whole panel below it
// (every workgroup factors the block itself in LDS and solves its own 64
// panel rows, so no workgroup waits for another).  Per window: one GEMM
// updates the next window's columns, and one rank-512 SYRK (an MFMA GEMM
// restricted to tiles on or below the diagonal, the reference's SYRK,
// mat.h:5471-5487) updates the rest on a second stream while the next
// window runs.  The low-precision drivers keep a recursion (potrf_rec):
// diagonal block, panel solve by the triangular-solve kernels, trailing
// GEMM.  The first failing column is recorded in a device flag that later
// launches check, so work stops at the same column the reference stops at.
#include <mutex>

#include "ops.hpp"

namespace mg {

constexpr int CHOL_NB = 64;
// columns per lane in the diagonal-block kernel (measured at n = 8192:
// 8 beats 16 by ~2%, 32 is 5% slower)
constexpr int CHOL_CS = 8;

// One workgroup factors the (at most 64 x 64) diagonal block, right-looking,
// with G = 64/CS lanes per row: thread t holds row t/G, columns CS*(t%G) ..
// +CS-1 in registers, so a step costs CS FMAs per lane (one lane per row
// would do up to 64).  Step J: the lanes holding column J write it,
// unscaled and with zeros above the diagonal, to LDS (by step parity); one
// barrier; every lane reads the pivot and the entries it needs and scales
// them itself, then updates its 16 columns unconditionally (entries above
// the diagonal take updates too and are never read); the lane holding
// (r, J) then keeps the finished L(r, J).  The loop runs over the G column
// slices, unrolled over the CS steps of one (static register indices, a
// small loop body).
template <typename T, int CS>
__global__ __launch_bounds__(CHOL_NB * (CHOL_NB / CS)) void chol_split_kernel(int n, T *A, long lda,
                                                                             int j0, int *info) {
  constexpr int NB = CHOL_NB, G = NB / CS;  // lanes per row
  __shared__ __align__(16) T col[2][NB];
  if (*(volatile int *)info != 0) return;  // an earlier column already failed
  const int t = threadIdx.x, r = t / G, g = t % G, c0 = CS * g;
  if (t < 2 * NB) col[t / NB][t % NB] = T(0);  // rows past n stay zero
  __syncthreads();
  T a[CS];
#pragma unroll
  for (int i = 0; i < CS; ++i) {
    const int c = c0 + i;
    a[i] = (r < n && c < n && c <= r) ? A[r + (long)c * lda] : T(0);
  }
#pragma unroll 1
  for (int gq = 0; gq < NB / CS; ++gq) {
#pragma unroll
    for (int sl = 0; sl < CS; ++sl) {
      const int J = CS * gq + sl;
      if (J >= n) break;
      T *cb = col[J & 1];
      if (g == gq && r < n) cb[r] = r >= J ? a[sl] : T(0);
      __syncthreads();
      const T d = cb[J];
      if (d <= T(0)) {  // the reference's test (every lane reads the same d); NaN passes it
        if (t == 0) *info = j0 + J + 1;
        return;
      }
      T l, inv;
      if constexpr (sizeof(T) == 4) {
        if (d >= 1e-30f) {
          inv = __builtin_amdgcn_rsqf(d);
          l = d * inv;
        } else {
          l = sqrtf(d);
          inv = 1.0f / l;
        }
      } else {
        l = sqrt(d);
        inv = T(1) / l;
      }
      const T lr = r > J ? cb[r] * inv : T(0);  // L(r, J) below the diagonal
      const T f = lr * inv;
      T cv[CS];
#pragma unroll
      for (int k = 0; k < CS; k += 16 / (int)sizeof(T)) {
        if constexpr (sizeof(T) == 4) {
          const float4 v = *reinterpret_cast<const float4 *>(cb + c0 + k);
          cv[k] = v.x, cv[k + 1] = v.y, cv[k + 2] = v.z, cv[k + 3] = v.w;
        } else {
          const double2 v = *reinterpret_cast<const double2 *>(cb + c0 + k);
          cv[k] = v.x, cv[k + 1] = v.y;
        }
      }
      // a(r, c) -= L(r, J) L(c, J), L(c, J) = cv[c] / l (zero for c < J)
#pragma unroll
      for (int i = 0; i < CS; ++i) a[i] = fma(-f, cv[i], a[i]);
      if (g == gq) a[sl] = r > J ? lr : (r == J ? l : a[sl]);
    }
  }
#pragma unroll
  for (int i = 0; i < CS; ++i) {
    const int c = c0 + i;
    if (r < n && c < n && c <= r) A[r + (long)c * lda] = a[i];
  }
}

// The same diagonal block, blocked, optionally with the panel below it:
// workgroup b holds the 64 x 64 block and, when R > 0, panel rows
// 64 + b R .. 64 + b R + R - 1 (relative to the block) in LDS (row-major,
// stride 68), and factors the block in sub-panels of SB columns, each
// workgroup redundantly, so that it can solve its own panel rows with no
// wait on another workgroup.  Per sub-panel:
//   1. every lane reads the SB x SB diagonal tile into its registers and
//      factors it there (each lane the whole tile, no cross-lane traffic:
//      bound by FMA issue, not by a chain of exchanges), and one lane per
//      row below the tile (block rows, then panel rows) solves its row,
//      x L11^T = a, from the same registers (x_J *= 1/L(J, J), then
//      x_c -= x_J L(c, J), the elimination's order);
//   2. lane 0 stores the tile; the rows below take the rank-SB update of
//      the block's trailing lower triangle (m x m) and of the panel rows'
//      trailing columns, 2 x 2 elements per thread;
// two workgroup barriers per SB columns.  Every workgroup writes its panel
// rows, the last one to finish the block.  A panel needs the full 64-column
// block (n = 64).  -1 at the first d <= 0 (the reference's test; NaN passes
// it) as the other kernels: nothing is written back.
constexpr int CB_LD = CHOL_NB + 4;

// SB consecutive values from 16-byte aligned LDS into registers
template <typename T, int SB>
__device__ __forceinline__ void lds_row(const T *p, T *x) {
#pragma unroll
  for (int c = 0; c < SB; c += 16 / (int)sizeof(T)) {
    if constexpr (sizeof(T) == 4) {
      const float4 v = *reinterpret_cast<const float4 *>(p + c);
      x[c] = v.x, x[c + 1] = v.y, x[c + 2] = v.z, x[c + 3] = v.w;
    } else {
      const double2 v = *reinterpret_cast<const double2 *>(p + c);
      x[c] = v.x, x[c + 1] = v.y;
    }
  }
}

// (e, count of 2 x 2 blocks of a lower triangle with nb block rows) ->
// block row and column, bc <= bi, row by row
__device__ __forceinline__ void tri_block(int e, int &bi, int &bc) {
  bi = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  while ((bi + 1) * (bi + 2) / 2 <= e) ++bi;
  while (bi * (bi + 1) / 2 > e) --bi;
  bc = e - bi * (bi + 1) / 2;
}

// a(R0.., C0..) -= x(R0..) x(C0..)^T over the SB columns at jb, for the
// 2 x 2 elements at rows R0, R0 + 1 and columns C0, C0 + 1 of S; w01, w10,
// w11 say which of the other three are written
template <typename T, int SB>
__device__ __forceinline__ void upd22(T *S, int jb, int R0, int C0, bool w01, bool w10, bool w11) {
  constexpr int LD = CB_LD;
  T xi0[SB], xi1[SB], xc0[SB], xc1[SB];
  lds_row<T, SB>(S + R0 * LD + jb, xi0);
  lds_row<T, SB>(S + (R0 + 1) * LD + jb, xi1);
  lds_row<T, SB>(S + C0 * LD + jb, xc0);
  lds_row<T, SB>(S + (C0 + 1) * LD + jb, xc1);
  T a00 = S[R0 * LD + C0], a01 = S[R0 * LD + C0 + 1];
  T a10 = S[(R0 + 1) * LD + C0], a11 = S[(R0 + 1) * LD + C0 + 1];
#pragma unroll
  for (int k = 0; k < SB; ++k) {
    a00 = fma(-xi0[k], xc0[k], a00);
    a01 = fma(-xi0[k], xc1[k], a01);
    a10 = fma(-xi1[k], xc0[k], a10);
    a11 = fma(-xi1[k], xc1[k], a11);
  }
  S[R0 * LD + C0] = a00;
  if (w01) S[R0 * LD + C0 + 1] = a01;
  if (w10) S[(R0 + 1) * LD + C0] = a10;
  if (w11) S[(R0 + 1) * LD + C0 + 1] = a11;
}

template <typename T, int SB, int R>
__global__ __launch_bounds__(256) void chol_blk_kernel(int n, int mrows, T *A, long lda, int j0,
                                                       int *info, unsigned *done) {
  constexpr int NB = CHOL_NB, LD = CB_LD, NR = NB + R;
  static_assert(R % 64 == 0 && R <= 192, "panel rows per workgroup");
  __shared__ __align__(16) T S[NR * LD];
  __shared__ int s_fail, s_last;
  const int t = threadIdx.x;
  // With a panel, every workgroup reads the block from A, so the factored
  // block may only be written once all have read it: each workgroup counts
  // itself out in *done (zero at launch) on every exit path, and the last
  // one (which holds the same factored block) writes it and resets *done
  // for the next launch on the stream.  Exits before the end are failures,
  // which write no block.
  auto last_out = [&]() -> bool {
    if constexpr (R == 0) {
      return true;
    } else {
      __syncthreads();
      if (t == 0) {
        const bool last = atomicAdd(done, 1u) == gridDim.x - 1;
        if (last) __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = last;
      }
      __syncthreads();
      return s_last != 0;
    }
  };
  const int pr0 = NB + blockIdx.x * R;                 // first panel row (relative)
  const int Rb = R > 0 ? max(0, min(R, mrows - blockIdx.x * R)) : 0;  // panel rows held
  {
    // thread t reads row t % 64 of columns t / 64 + 4k (coalesced along the
    // column); elements outside the lower triangle (or past the panel) read
    // A(0, 0) instead and are replaced by zero, so the loads need no branch
    constexpr int PER = NB * NB / 256, PPER = R * NB / 256;
    const int r = t % NB, c0 = t / NB;
    const T *p = A + r + (long)c0 * lda;
    const long step = 4 * lda;
    T v[PER], pv[PPER > 0 ? PPER : 1];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = c0 + 4 * k;
      const bool in = r < n && c < n && c <= r;
      const T x = *(in ? p + k * step : A);
      v[k] = in ? x : T(0);
    }
    // panel: chunk k / 16 of 64 rows, column c0 + 4 (k % 16)
#pragma unroll
    for (int k = 0; k < PPER; ++k) {
      const int i = 64 * (k / 16) + r, c = c0 + 4 * (k % 16);
      const bool in = i < Rb;
      const T x = *(in ? A + (pr0 + i) + (long)c * lda : A);
      pv[k] = in ? x : T(0);
    }
    // an earlier column already failed?  (With a panel, workgroup 0 of
    // this launch may be writing the flag: one read per workgroup, so that
    // the whole workgroup takes the same branch.)
    if (t == 0) s_fail = *(volatile int *)info;
    __syncthreads();
    if (s_fail != 0) {
      last_out();
      return;
    }
    __syncthreads();
    if (t == 0) s_fail = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) S[r * LD + c0 + 4 * k] = v[k];
#pragma unroll
    for (int k = 0; k < PPER; ++k) S[(NB + 64 * (k / 16) + r) * LD + c0 + 4 * (k % 16)] = pv[k];
  }
  __syncthreads();
  for (int jb = 0; jb < n; jb += SB) {
    const int w = min(SB, n - jb), m = max(0, n - jb - SB);
    const int below = m + Rb;  // rows below the tile in this workgroup
    // 1. the tile in every lane (lower triangle used)
    T a[SB][SB], inv[SB];
#pragma unroll
    for (int i = 0; i < SB; ++i) lds_row<T, SB>(S + (jb + i) * LD + jb, a[i]);
    // straight-line: the pivot test only records the first failing column
    // (later values are garbage and are never written), columns past n
    // (J >= w, a short last block) get l = 1/l = 0, which leaves the zero
    // rows and columns past n unchanged
    int fail = 0;
#pragma unroll
    for (int J = 0; J < SB; ++J) {
      const bool live = J < w;
      const T d = a[J][J];
      fail = (live && fail == 0 && d <= T(0)) ? jb + J + 1 : fail;
      T l, iv;
      if constexpr (sizeof(T) == 4) {
        // 1/sqrt on the hardware estimate; pivots below 1e-30 are scaled by
        // 2^100 first (exact), so the estimate never sees a denormal
        const bool tiny = d < 1e-30f;
        iv = __builtin_amdgcn_rsqf(tiny ? d * 0x1p100f : d);
        iv = tiny ? iv * 0x1p50f : iv;
        l = d * iv;
      } else {
        l = sqrt(d);
        iv = T(1) / l;
      }
      a[J][J] = live ? l : T(0);
      inv[J] = iv = live ? iv : T(0);
#pragma unroll
      for (int i = J + 1; i < SB; ++i) a[i][J] *= iv;
#pragma unroll
      for (int c = J + 1; c < SB; ++c)
#pragma unroll
        for (int i = c; i < SB; ++i) a[i][c] = fma(-a[i][J], a[c][J], a[i][c]);
    }
    if (fail) {
      if (t == 0) s_fail = fail;
    } else if (t < below) {
      T *row = S + (t < m ? jb + SB + t : NB + (t - m)) * LD + jb;
      T x[SB];
      lds_row<T, SB>(row, x);
#pragma unroll
      for (int J = 0; J < SB; ++J) {
        x[J] *= inv[J];
#pragma unroll
        for (int c = J + 1; c < SB; ++c) x[c] = fma(-x[J], a[c][J], x[c]);
      }
#pragma unroll
      for (int c = 0; c < SB; ++c) row[c] = x[c];
    }
    __syncthreads();
    if (s_fail) {
      if (t == 0 && blockIdx.x == 0) *info = j0 + s_fail;
      last_out();
      return;
    }
    // 2. every lane has read the tile: lane 0 stores it (rows below and the
    // trailing part are not in it)
    if (t == 0) {
#pragma unroll
      for (int i = 0; i < SB; ++i)
#pragma unroll
        for (int c = 0; c < SB; ++c) S[(jb + i) * LD + jb + c] = c <= i ? a[i][c] : T(0);
    }
    if (m > 0) {
      // 2 x 2 blocks: the block's trailing triangle (block (bi, bc), bc <= bi,
      // rows 2bi, 2bi + 1, columns 2bc, 2bc + 1 past the tile; a row or
      // column at m lies inside the 64-row block, as m is odd only when
      // n < 64, and is never written), then the panel rows x the same
      // columns (m is even there: n = 64)
      const int nb = (m + 1) / 2, ntri = nb * (nb + 1) / 2;
      const int nblk = ntri + ((Rb + 1) / 2) * nb;
      for (int e = t; e < nblk; e += 256) {
        if (e < ntri) {
          int bi, bc;
          tri_block(e, bi, bc);
          const int i0 = 2 * bi, c0 = 2 * bc;
          upd22<T, SB>(S, jb, jb + SB + i0, jb + SB + c0, c0 + 1 <= i0, i0 + 1 < m,
                       i0 + 1 < m && c0 + 1 < m);
        } else {
          const int q = e - ntri, bi = q / nb, bc = q % nb;
          const int i0 = 2 * bi;
          upd22<T, SB>(S, jb, NB + i0, jb + SB + 2 * bc, true, i0 + 1 < Rb, i0 + 1 < Rb);
        }
      }
    }
    __syncthreads();
  }
  if constexpr (R > 0) {
    for (int e = t; e < Rb * NB; e += 256) {
      const int i = e % Rb, c = e / Rb;
      A[(pr0 + i) + (long)c * lda] = S[(NB + i) * LD + c];
    }
  }
  if (last_out()) {
    for (int e = t; e < NB * NB; e += 256) {
      const int r = e % NB, c = e / NB;
      if (r < n && c < n && c <= r) A[r + (long)c * lda] = S[r * LD + c];
    }
  }
}

// Diagonal-block kernel: 1 = blocked with 8-column sub-panels (default),
// 0 = eight lanes per row (chol_split_kernel); MATGPU_CHOL_LEAF overrides.
// Measured on MI355X for one 64 x 64 block alone: fp32 17.9 / 19.4 us,
// fp64 24.8 / 28.4 us; 16-column sub-panels took 21.8 (fp32) and 39.8 us
// (fp64): a 16 x 16 tile held whole in every lane is ~1100 instructions of
// one wavefront, and spills in fp64.
static int chol_leaf_kind() {
  static const int v = [] {
    const char *e = getenv("MATGPU_CHOL_LEAF");
    return e ? atoi(e) : 1;
  }();
  return v;
}

template <typename T>
static void potrf_rec(hipStream_t s, int n, T *A, long lda, int j0, int *info, Mul mul) {
  if (n <= CHOL_NB) {
    KScope ks(s, "chol_leaf", (double)n * n * sizeof(T));
    if (chol_leaf_kind() == 1)
      hipLaunchKernelGGL((chol_blk_kernel<T, 8, 0>), dim3(1), dim3(256), 0, s, n, 0, A, lda, j0, info,
                         nullptr);
    else
      hipLaunchKernelGGL((chol_split_kernel<T, CHOL_CS>), dim3(1),
                         dim3(CHOL_NB * (CHOL_NB / CHOL_CS)), 0, s, n, A, lda, j0, info);
    MG_HIP(hipGetLastError());
    return;
  }
  int h = ((n / 2 + CHOL_NB - 1) / CHOL_NB) * CHOL_NB;
  if (h >= n) h = n - CHOL_NB;
  const int m = n - h;
  potrf_rec<T>(s, h, A, lda, j0, info, mul);
  T *A21 = A + h;
  T *A22 = A + h + (long)h * lda;
  // A21 := A21 L11^-T  <=>  L11 * A21^T = A21^T (A21^T is row-major in A21's storage).
  // The panel stays near fp32 accuracy (fp8 only ever enters the trailing
  // update A22), so an fp8 request runs the solve with 3-pass bf16 splits.
  const Mul tmul = mul == Mul::Fp8 ? Mul::Bf16x3 : mul;
  trsm_left<T>(s, true, false, h, m, View<const T>{A, 1, lda}, View<T>{A21, lda, 1}, tmul);
  // A22 -= A21 A21^T on the lower triangle
  View<const T> a21{A21, 1, lda};
  const bool big = (double)m * m * h >= kLowpMinWork;
  if constexpr (std::is_same<T, float>::value) {
    if (!big) {
      gemm<T>(s, m, m, h, T(-1), a21, a21.t(), T(1), View<T>{A22, 1, lda}, true);
    } else if (mul == Mul::Fp8) {
      gemm_fp8(s, m, m, h, -1.0f, a21, a21.t(), View<float>{A22, 1, lda}, true);
    } else if (mul != Mul::Native) {
      gemm_bf16x(s, m, m, h, -1.0f, a21, a21.t(), View<float>{A22, 1, lda}, (int)mul, true);
    } else {
      gemm<T>(s, m, m, h, T(-1), a21, a21.t(), T(1), View<T>{A22, 1, lda}, true);
    }
  } else {
    gemm<T>(s, m, m, h, T(-1), a21, a21.t(), T(1), View<T>{A22, 1, lda}, true);
  }
  potrf_rec<T>(s, m, A22, lda, j0 + h, info, mul);
}

// A -= a a^T on the lower triangle of the m x m matrix C (a: m x k), in
// the precision the recursion uses for its trailing updates.
template <typename T>
static void syrk_lower(hipStream_t s, long m, long k, View<const T> a, T *C, long lda, Mul mul,
                       int grid_cap = 0) {
  if (m <= 0 || k <= 0) return;
  if constexpr (std::is_same<T, float>::value) {
    if ((double)m * m * k >= kLowpMinWork) {
      if (mul == Mul::Fp8) {
        gemm_fp8(s, m, m, k, -1.0f, a, a.t(), View<float>{C, 1, lda}, true);
        return;
      }
      if (mul != Mul::Native) {
        gemm_bf16x(s, m, m, k, -1.0f, a, a.t(), View<float>{C, 1, lda}, (int)mul, true);
        return;
      }
    }
  }
  gemm<T>(s, m, m, k, T(-1), a, a.t(), T(1), View<T>{C, 1, lda}, true, grid_cap);
}

// Right-looking blocked Cholesky with look-ahead over NB-wide column
// blocks (the same two-stream schedule as the LU's):
//   caller's stream: block k's diagonal block (recursive, as above), then
//                    its panel A21 := A21 L11^-T;
//   side stream:     block k's update of block k+1's columns (the only
//                    part the caller's next step waits for), then of
//                    everything right of them.
// The caller's stream factors block k+1 while the side stream runs block
// k's large trailing update; the two never touch the same columns at the
// same time, and updates of one column block arrive in block order.
template <typename T>
static void potrf_lookahead(hipStream_t s, int n, T *A, long lda, int *info, Mul mul, int NB) {
  std::lock_guard<std::mutex> lk(lookahead_mutex());
  hipStream_t u = side_stream();
  const int nblk = (n + NB - 1) / NB;
  hipEvent_t start = pooled_event(2 * nblk);
  MG_HIP(hipEventRecord(start, s));
  MG_HIP(hipStreamWaitEvent(u, start, 0));
  const Mul tmul = mul == Mul::Fp8 ? Mul::Bf16x3 : mul;
  for (int k = 0; k < nblk; ++k) {
    const int b0 = k * NB, b1 = std::min(n, b0 + NB), nbk = b1 - b0, m = n - b1;
    if (k >= 1) MG_HIP(hipStreamWaitEvent(s, pooled_event(2 * (k - 1) + 1), 0));
    T *Akk = A + b0 + (long)b0 * lda;
    potrf_rec<T>(s, nbk, Akk, lda, b0, info, mul);
    if (m == 0) break;
    T *A21 = A + b1 + (long)b0 * lda;
    trsm_left<T>(s, true, false, nbk, m, View<const T>{Akk, 1, lda}, View<T>{A21, lda, 1}, tmul);
    hipEvent_t evP = pooled_event(2 * k), evU = pooled_event(2 * k + 1);
    MG_HIP(hipEventRecord(evP, s));
    MG_HIP(hipStreamWaitEvent(u, evP, 0));
    const int c2 = std::min(n, b1 + NB), w2 = c2 - b1;
    const View<const T> a21{A21, 1, lda};
    // block k+1's columns (rows b1..n-1), lower part of its diagonal block
    gemm<T>(u, m, w2, nbk, T(-1), a21, a21.t(), T(1), View<T>{A + b1 + (long)b1 * lda, 1, lda},
            true);
    MG_HIP(hipEventRecord(evU, u));
    // the rest of the trailing matrix
    syrk_lower<T>(u, n - c2, nbk, View<const T>{A21 + w2, 1, lda}, A + c2 + (long)c2 * lda, lda,
                  mul);
  }
  hipEvent_t done = pooled_event(2 * nblk + 1);
  MG_HIP(hipEventRecord(done, u));
  MG_HIP(hipStreamWaitEvent(s, done, 0));
}

// The 64-column block at A (n <= 64 columns) and the mrows panel rows
// below it in one launch (chol_blk_kernel with R panel rows per workgroup):
// block factored, panel := panel L^-T.
// panel rows per workgroup for fp32 (64, 128 or 192; MATGPU_CHOL_R), fp64
// always 64 (LDS)
static int chol_panel_rows() {
  static const int v = [] {
    const char *e = getenv("MATGPU_CHOL_R");
    return e ? atoi(e) : 64;
  }();
  return v;
}
template <typename T>
static void chol_panel(hipStream_t s, int n, int mrows, T *A, long lda, int j0, int *info) {
  MG_CHECK(mrows == 0 || n == CHOL_NB, "chol_panel: a panel needs a full 64-column block");
  const int R = sizeof(T) == 4 ? chol_panel_rows() : 64;
  const int grid = mrows > 0 ? (int)ceil_div(mrows, R) : 1;
  bool fresh = false;
  unsigned *done = (unsigned *)workspace(s).get("chol.done", 4, &fresh);
  if (fresh) fill_u32(s, done, 1, 0u);
  KScope ks(s, "chol_panel", ((double)n * n + 2.0 * mrows * n) * sizeof(T));
  if (R == 192)
    hipLaunchKernelGGL((chol_blk_kernel<T, 8, 192>), dim3(grid), dim3(256), 0, s, n, mrows, A, lda,
                       j0, info, done);
  else if (R == 128)
    hipLaunchKernelGGL((chol_blk_kernel<T, 8, 128>), dim3(grid), dim3(256), 0, s, n, mrows, A, lda,
                       j0, info, done);
  else
    hipLaunchKernelGGL((chol_blk_kernel<T, 8, 64>), dim3(grid), dim3(256), 0, s, n, mrows, A, lda,
                       j0, info, done);
  MG_HIP(hipGetLastError());
}

// Workgroups of the side-stream SYRK: one per CU by default, each walking
// several tiles, so the SYRK (which has slack: it runs beside a whole
// window of panels) shares the CUs less with the critical path; 0 = one
// workgroup per tile; MATGPU_SYRK_CAP overrides.  Measured, fp32 n = 8192:
// 8.16 ms per solve capped at 256 workgroups, 8.31 uncapped, 9.04 at 128.
static int syrk_cap() {
  static const int v = [] {
    const char *e = getenv("MATGPU_SYRK_CAP");
    return e ? atoi(e) : -1;
  }();
  return v >= 0 ? v : num_cus();
}

// Windowed Cholesky: W-column windows of 64-column blocks.
//   caller's stream, per block k of window [w0, w1):
//     block k's columns (rows bk..n-1) -= the window's earlier panels
//     (left-looking inside the window, one GEMM, K = bk - w0 < W);
//     block k and its whole panel in one launch (chol_panel);
//   after the window: the next window's columns -= this window's panels
//     (one GEMM, K = W);
//   side stream: everything right of the next window -= this window's
//     panels (one rank-W SYRK), overlapping the next window.
// The next window's update waits for the previous window's side-stream
// SYRK (both write the columns right of this window), so one SYRK is in
// flight at a time and updates of a column arrive in window order.
template <typename T>
static void potrf_window(hipStream_t s, int n, T *A, long lda, int *info, int W) {
  std::lock_guard<std::mutex> lk(lookahead_mutex());
  hipStream_t u = side_stream();
  const int nwin = (n + W - 1) / W;
  hipEvent_t start = pooled_event(2 * nwin);
  MG_HIP(hipEventRecord(start, s));
  MG_HIP(hipStreamWaitEvent(u, start, 0));
  auto at = [&](long r, long c) { return A + r + c * lda; };
  bool side = false;
  for (int j = 0; j < nwin; ++j) {
    const int w0 = j * W, w1 = std::min(n, w0 + W);
    for (int bk = w0; bk < w1; bk += CHOL_NB) {
      const int nbk = std::min(CHOL_NB, n - bk), mrows = n - bk - nbk;
      if (bk > w0) {
        const View<const T> p{at(bk, w0), 1, lda};
        gemm<T>(s, n - bk, nbk, bk - w0, T(-1), p, p.t(), T(1), View<T>{at(bk, bk), 1, lda}, true);
      }
      chol_panel<T>(s, nbk, mrows, at(bk, bk), lda, bk, info);
    }
    if (w1 >= n) break;
    const int w2 = std::min(n, w1 + W);
    if (j >= 1 && side) MG_HIP(hipStreamWaitEvent(s, pooled_event(2 * (j - 1) + 1), 0));
    const View<const T> pn{at(w1, w0), 1, lda};
    gemm<T>(s, n - w1, w2 - w1, w1 - w0, T(-1), pn, pn.t(), T(1), View<T>{at(w1, w1), 1, lda}, true);
    side = false;
    if (w2 < n) {
      hipEvent_t evP = pooled_event(2 * j), evR = pooled_event(2 * j + 1);
      MG_HIP(hipEventRecord(evP, s));
      MG_HIP(hipStreamWaitEvent(u, evP, 0));
      syrk_lower<T>(u, n - w2, w1 - w0, View<const T>{at(w2, w0), 1, lda}, at(w2, w2), lda,
                    Mul::Native, syrk_cap());
      MG_HIP(hipEventRecord(evR, u));
      side = true;
    }
  }
  hipEvent_t done = pooled_event(2 * nwin + 1);
  MG_HIP(hipEventRecord(done, u));
  MG_HIP(hipStreamWaitEvent(s, done, 0));
}

// Cholesky driver: 0 = recursion / look-ahead as below, 1 = windowed
// (default); MATGPU_CHOL_WIN overrides, MATGPU_CHOL_W sets the window
// (default 512).  Measured on MI355X, fp32 n = 8192 with 8 right-hand
// sides, whole solve: look-ahead 12.7 ms; windowed W = 384 / 512 / 768:
// 9.38 / 9.34 / 9.46 ms (64 panel rows per workgroup; 128 and 192 rows:
// 10.2 and 10.8 ms at W = 512, the redundant diagonal block is cheap next
// to a workgroup's panel rows).
static int chol_win_mode() {
  static const int v = [] {
    const char *e = getenv("MATGPU_CHOL_WIN");
    return e ? atoi(e) : 1;
  }();
  return v;
}
static int chol_win_width() {
  static const int v = [] {
    const char *e = getenv("MATGPU_CHOL_W");
    const int w = e ? atoi(e) : 512;
    return w >= CHOL_NB && w % CHOL_NB == 0 ? w : 512;
  }();
  return v;
}

template <typename T>
void potrf_lower(hipStream_t s, int n, T *A, long lda, int *info, Mul mul) {
  fill_int(s, info, 1, 0);
  const int NB = lookahead_block();
  // Look-ahead pays while the factorization is latency-bound.  Measured on
  // MI355X (fp32, NB = 256, against the recursion): +8% at n = 4096, +14% at
  // 8192, +17% at 12288, even at 16384; fp64 at 16384 is 8% slower, and at
  // 32768 both precisions lose to the recursion's few large-K products
  // (the blocked update re-reads the trailing matrix n/NB times).  Low-
  // precision updates keep the recursion: n/NB separately rounded FP8
  // updates need more refinement steps.  MATGPU_CHOL_LA_MAX overrides.
  static const int la_env = [] {
    const char *e = getenv("MATGPU_CHOL_LA_MAX");
    return e ? atoi(e) : -1;
  }();
  const int la_max = la_env >= 0 ? la_env : (sizeof(T) == 4 ? 16384 : 8192);
  if (mul == Mul::Native && chol_win_mode() == 1 && n >= 2 * chol_win_width() && n <= la_max)
    potrf_window<T>(s, n, A, lda, info, chol_win_width());
  else if (mul == Mul::Native && n >= 4 * NB && n <= la_max && NB >= CHOL_NB && NB % CHOL_NB == 0)
    potrf_lookahead<T>(s, n, A, lda, info, mul, NB);
  else
    potrf_rec<T>(s, n, A, lda, 0, info, mul);
}

template void potrf_lower<float>(hipStream_t, int, float *, long, int *, Mul);
template void potrf_lower<double>(hipStream_t, int, double *, long, int *, Mul);

}  // namespace mg
