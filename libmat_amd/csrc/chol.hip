// Lower Cholesky factorization A = L L^T (mat_chol, mat.h:6062-6095).
//
// Reference semantics (x86 path mat__chol_scalar_generic_, mat.h:6031-6060):
// only the lower triangle of A is read; column j's pivot is
// d = a(j,j) - sum_k l(j,k)^2, the factorization fails (-1) at the first
// d <= 0, l(j,j) = sqrt(d) and the column below is multiplied by 1/l(j,j).
//
// MI355X design: recursive blocking.  A 64-column diagonal block is
// factored by one workgroup in registers (right-looking: scale the column,
// rank-1 update of the trailing columns); the panel below it is solved by
// the triangular-solve kernels (A21 := A21 L11^-T), and the trailing lower
// triangle is updated by one MFMA GEMM restricted to tiles on or below the
// diagonal (the reference's SYRK, mat.h:5471-5487).  The first failing
// column is recorded in a device flag that later leaves check, so work stops
// at the same column the reference stops at.
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

// The same diagonal block, blocked: the 64 x 64 block in LDS (row-major,
// stride 68), factored in sub-panels of SB columns; per sub-panel
//   1. wavefront 0 reads the SB x SB diagonal tile into every lane's
//      registers and factors it there (each lane the whole tile, no
//      cross-lane traffic: bound by FMA issue, not by a chain of exchanges),
//      lane 0 stores it, and lanes 0 .. m-1 solve their row below the tile,
//      x L11^T = a, from the same registers (x_J *= 1/L(J, J), then
//      x_c -= x_J L(c, J), the elimination's order);
//   2. the trailing lower triangle (m x m) takes the rank-SB update, 2 x 2
//      elements per thread;
// two workgroup barriers per SB columns.  -1 at the first d <= 0 (the
// reference's test; NaN passes it) as the other kernels: nothing is written
// back.
constexpr int CB_LD = CHOL_NB + 4;

// v of the first active lane, as a wavefront-uniform value
template <typename T>
__device__ __forceinline__ T first_lane(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}

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

template <typename T, int SB>
__global__ __launch_bounds__(256) void chol_blk_kernel(int n, T *A, long lda, int j0, int *info) {
  constexpr int NB = CHOL_NB, LD = CB_LD;
  __shared__ __align__(16) T S[NB * LD];
  __shared__ int s_fail;
  const int t = threadIdx.x;
  const int failed = *(volatile int *)info;  // an earlier column already failed
  {
    // thread t reads row t % 64 of columns t / 64 + 4k (coalesced along the
    // column); elements outside the lower triangle read A(0, 0) instead and
    // are replaced by zero, so the loads need no branch
    constexpr int PER = NB * NB / 256;
    const int r = t % NB, c0 = t / NB;
    const T *p = A + r + (long)c0 * lda;
    const long step = 4 * lda;
    T v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = c0 + 4 * k;
      const bool in = r < n && c < n && c <= r;
      const T x = *(in ? p + k * step : A);
      v[k] = in ? x : T(0);
    }
    if (failed != 0) return;
    if (t == 0) s_fail = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) S[r * LD + c0 + 4 * k] = v[k];
  }
  __syncthreads();
  for (int jb = 0; jb < n; jb += SB) {
    const int w = min(SB, n - jb), m = n - jb - SB;
    if (t < 64) {
      // 1. the tile in every lane of wavefront 0 (lower triangle used)
      T a[SB][SB], inv[SB];
#pragma unroll
      for (int i = 0; i < SB; ++i) lds_row<T, SB>(S + (jb + i) * LD + jb, a[i]);
      int fail = 0;
#pragma unroll
      for (int J = 0; J < SB; ++J) {
        inv[J] = T(0);  // columns past n: no scaling, no update
        if (J < w && fail == 0) {
          const T d = first_lane(a[J][J]);
          if (d <= T(0)) {
            fail = jb + J + 1;
          } else {
            T l, iv;
            if constexpr (sizeof(T) == 4) {
              if (d >= 1e-30f) {
                iv = __builtin_amdgcn_rsqf(d);
                l = d * iv;
              } else {
                l = sqrtf(d);
                iv = 1.0f / l;
              }
            } else {
              l = sqrt(d);
              iv = T(1) / l;
            }
            a[J][J] = l;
            inv[J] = iv;
#pragma unroll
            for (int i = J + 1; i < SB; ++i) a[i][J] *= iv;
#pragma unroll
            for (int c = J + 1; c < SB; ++c)
#pragma unroll
              for (int i = c; i < SB; ++i) a[i][c] = fma(-a[i][J], a[c][J], a[i][c]);
          }
        }
      }
      if (fail) {
        if (t == 0) s_fail = fail;
      } else {
        // the wavefront has read the whole tile: its LDS ops run in order,
        // so lane 0 may overwrite it now
        if (t == 0) {
#pragma unroll
          for (int i = 0; i < SB; ++i)
#pragma unroll
            for (int c = 0; c < SB; ++c) S[(jb + i) * LD + jb + c] = c <= i ? a[i][c] : T(0);
        }
        if (t < m) {
          T *row = S + (jb + SB + t) * LD + jb;
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
      }
    }
    __syncthreads();
    if (s_fail) {
      if (t == 0) *info = j0 + s_fail;
      return;
    }
    if (m > 0) {
      // 2. rank-SB update of the trailing m x m lower triangle in 2 x 2
      // blocks; block (bi, bc), bc <= bi, rows 2bi, 2bi + 1, columns
      // 2bc, 2bc + 1 (a row or column at m lies inside the 64-row block, as
      // m is odd only when n < 64, and is never written)
      const int nb = (m + 1) / 2, nblk = nb * (nb + 1) / 2;
      for (int e = t; e < nblk; e += 256) {
        int bi = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
        while ((bi + 1) * (bi + 2) / 2 <= e) ++bi;
        while (bi * (bi + 1) / 2 > e) --bi;
        const int bc = e - bi * (bi + 1) / 2;
        const int i0 = 2 * bi, c0 = 2 * bc;
        const int R0 = jb + SB + i0, C0 = jb + SB + c0;
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
        if (c0 + 1 <= i0) S[R0 * LD + C0 + 1] = a01;
        if (i0 + 1 < m) {
          S[(R0 + 1) * LD + C0] = a10;
          if (c0 + 1 < m) S[(R0 + 1) * LD + C0 + 1] = a11;
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e % NB, c = e / NB;
    if (r < n && c < n && c <= r) A[r + (long)c * lda] = S[r * LD + c];
  }
}

// Diagonal-block kernel: 1 = blocked with 8-column sub-panels (default),
// 2 = blocked with 16-column sub-panels, 0 = eight lanes per row
// (chol_split_kernel); MATGPU_CHOL_LEAF overrides.  Measured on MI355X for
// one 64 x 64 block alone: fp32 17.9 / 21.8 / 19.4 us, fp64 24.8 / 39.8 /
// 28.4 us (a 16 x 16 tile held whole in every lane is ~1100 instructions
// of one wavefront, and spills in fp64).
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
      hipLaunchKernelGGL((chol_blk_kernel<T, 8>), dim3(1), dim3(256), 0, s, n, A, lda, j0, info);
    else if (chol_leaf_kind() == 2)
      hipLaunchKernelGGL((chol_blk_kernel<T, 16>), dim3(1), dim3(256), 0, s, n, A, lda, j0, info);
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
      gemm_bf16x(s, m, m, h, -1.0f, a21, a21.t(), View<float>{A22, 1, lda}, (int)mul);
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
static void syrk_lower(hipStream_t s, long m, long k, View<const T> a, T *C, long lda, Mul mul) {
  if (m <= 0 || k <= 0) return;
  if constexpr (std::is_same<T, float>::value) {
    if ((double)m * m * k >= kLowpMinWork) {
      if (mul == Mul::Fp8) {
        gemm_fp8(s, m, m, k, -1.0f, a, a.t(), View<float>{C, 1, lda}, true);
        return;
      }
      if (mul != Mul::Native) {
        gemm_bf16x(s, m, m, k, -1.0f, a, a.t(), View<float>{C, 1, lda}, (int)mul);
        return;
      }
    }
  }
  gemm<T>(s, m, m, k, T(-1), a, a.t(), T(1), View<T>{C, 1, lda}, true);
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
  if (mul == Mul::Native && n >= 4 * NB && n <= la_max && NB >= CHOL_NB && NB % CHOL_NB == 0)
    potrf_lookahead<T>(s, n, A, lda, info, mul, NB);
  else
    potrf_rec<T>(s, n, A, lda, 0, info, mul);
}

template void potrf_lower<float>(hipStream_t, int, float *, long, int *, Mul);
template void potrf_lower<double>(hipStream_t, int, double *, long, int *, Mul);

}  // namespace mg
