// Lower Cholesky factorization A = L L^T (mat_chol, mat.h:6062-6095).
//
// Reference semantics (x86 path mat__chol_scalar_generic_, mat.h:6031-6060):
// only the lower triangle of A is read; column j's pivot is
// d = a(j,j) - sum_k l(j,k)^2, the factorization fails (-1) at the first
// d <= 0, l(j,j) = sqrt(d) and the column below is multiplied by 1/l(j,j).
//
// MI355X design: recursive blocking.  A 64-column diagonal block is
// factored by one wavefront in registers (right-looking: scale the column,
// rank-1 update of the trailing lower triangle); the panel below it is solved by
// the triangular-solve kernels (A21 := A21 L11^-T), and the trailing lower
// triangle is updated by one MFMA GEMM restricted to tiles on or below the
// diagonal (the reference's SYRK, mat.h:5471-5487).  The first failing
// column is recorded in a device flag that later leaves check, so work stops
// at the same column the reference stops at.
#include "ops.hpp"

namespace mg {

constexpr int CHOL_NB = 64;

template <typename T>
__device__ __forceinline__ T lane_value(T v, int src) {
  if constexpr (sizeof(T) == 4) {
    return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), src));
  } else {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  }
}

// Column j of the one-wavefront leaf, then the columns after it; compile-time
// recursion so that every register index is static.
template <typename T, int J, int NB>
__device__ __forceinline__ void chol_steps(T (&a)[NB], int i, int n, int j0, int *info, bool &failed) {
  if constexpr (J < NB) {
    if (J >= n) return;
    const T d = lane_value(a[J], J);
    if (d <= T(0)) {  // the reference's test; NaN passes it
      if (i == 0) *info = j0 + J + 1;
      failed = true;
      return;
    }
    const T l = sqrt(d);
    const T inv = T(1) / l;
    a[J] = i == J ? l : (i > J ? a[J] * inv : a[J]);
    const T lij = a[J];
#pragma unroll
    for (int c = J + 1; c < NB; ++c) {
      const T lcj = lane_value(lij, c);
      a[c] = i >= c ? fma(-lij, lcj, a[c]) : a[c];
    }
    chol_steps<T, J + 1, NB>(a, i, n, j0, info, failed);
  }
}

// One wavefront factors the NB x NB diagonal block: lane i holds row i in
// registers; column j's entries are broadcast with v_readlane (static lane
// and register indices, no barrier); right-looking rank-1 updates.  The
// block goes through LDS only to be read and written column-coalesced.
template <typename T>
__global__ __launch_bounds__(64) void chol_leaf_kernel(int n, T *A, long lda, int j0, int *info) {
  constexpr int NB = CHOL_NB;
  __shared__ T S[NB * (NB + 1)];  // S[c*(NB+1) + r] = A(r, c)
  if (*(volatile int *)info != 0) return;  // an earlier column already failed
  const int i = threadIdx.x;
  for (int c = 0; c < NB; ++c) S[c * (NB + 1) + i] = (i < n && c < n && i >= c) ? A[i + (long)c * lda] : T(0);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  T a[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) a[c] = S[c * (NB + 1) + i];
  bool failed = false;
  chol_steps<T, 0, NB>(a, i, n, j0, info, failed);
  if (failed) return;
#pragma unroll
  for (int c = 0; c < NB; ++c) S[c * (NB + 1) + i] = a[c];
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  for (int c = 0; c < NB; ++c)
    if (i < n && c < n && i >= c) A[i + (long)c * lda] = S[c * (NB + 1) + i];
}

template <typename T>
static void potrf_rec(hipStream_t s, int n, T *A, long lda, int j0, int *info, Mul mul) {
  if (n <= CHOL_NB) {
    KScope ks(s, "chol_leaf", (double)n * n * sizeof(T));
    hipLaunchKernelGGL((chol_leaf_kernel<T>), dim3(1), dim3(64), 0, s, n, A, lda, j0, info);
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

template <typename T>
void potrf_lower(hipStream_t s, int n, T *A, long lda, int *info, Mul mul) {
  fill_int(s, info, 1, 0);
  potrf_rec<T>(s, n, A, lda, 0, info, mul);
}

template void potrf_lower<float>(hipStream_t, int, float *, long, int *, Mul);
template void potrf_lower<double>(hipStream_t, int, double *, long, int *, Mul);

}  // namespace mg
