// Left-side triangular solves B := op(T)^-1 B on strided views.
//
// Used for U12 = L11^-1 A12 inside the LU (the reference's TRSM-by-axpy,
// mat.h:4264-4271), for the Cholesky panel A21 L11^-T, and for the
// substitutions of mat_solve / mat_solve_spd / mat_solve_tri* (the scalar
// TRSVs, mat.h:4584-4803).
//
// Recursive blocking: split T at a multiple of 64, solve the leading block,
// update the other half of B with one GEMM (MFMA), solve the trailing
// block.  Leaves (n <= 64) run one thread per right-hand-side column with
// the column in registers and the 64x64 diagonal block in LDS, stored
// column-major so each elimination step reads a contiguous LDS run.
// Non-unit leaves divide by the diagonal, as the reference's TRSVs do.
#include <type_traits>

#include "ops.hpp"

namespace mg {

template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

constexpr int TRSM_TB = 64;

constexpr int TRSM_NT = 64;  // one wavefront per workgroup: more workgroups per solve

template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(TRSM_NT) void trsm_leaf_kernel(int n, int nrhs, View<const T> Tm,
                                                            View<T> B) {
  constexpr int TB = TRSM_TB;
  __shared__ T S[TB * (TB + 1)];  // S[c*(TB+1) + r] = T(r, c)
  {
    // issue every load before the LDS stores (one latency, not 64)
    constexpr int PER = TB * TB / TRSM_NT;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = threadIdx.x + u * TRSM_NT;
      const int r = Tm.rs == 1 ? idx % TB : idx / TB;
      const int c = Tm.rs == 1 ? idx / TB : idx % TB;
      v[u] = (r < n && c < n) ? Tm(r, c) : T(0);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = threadIdx.x + u * TRSM_NT;
      const int r = Tm.rs == 1 ? idx % TB : idx / TB;
      const int c = Tm.rs == 1 ? idx / TB : idx % TB;
      S[c * (TB + 1) + r] = v[u];
    }
  }
  __syncthreads();
  const int j = blockIdx.x * TRSM_NT + threadIdx.x;
  if (j >= nrhs) return;
  T x[TB];
  sfor<0, TB>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    x[r] = (r < n) ? B(r, j) : T(0);
  });
  if constexpr (LOWER) {
    sfor<0, TB>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      if (c < n) {
        if constexpr (!UNIT) x[c] = x[c] / S[c * (TB + 1) + c];
        const T xc = x[c];
#pragma unroll
        for (int r = c + 1; r < TB; ++r) x[r] = fma(-S[c * (TB + 1) + r], xc, x[r]);
      }
    });
  } else {
    sfor<0, TB>([&](auto cc) {
      constexpr int c = TB - 1 - decltype(cc)::value;
      if (c < n) {
        if constexpr (!UNIT) x[c] = x[c] / S[c * (TB + 1) + c];
        const T xc = x[c];
#pragma unroll
        for (int r = 0; r < c; ++r) x[r] = fma(-S[c * (TB + 1) + r], xc, x[r]);
      }
    });
  }
  sfor<0, TB>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    if (r < n) B(r, j) = x[r];
  });
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
  if (n <= TRSM_TB && nrhs <= 4096) {
    // small: one launch of the register-resident leaf
    dim3 grid(ceil_div(nrhs, TRSM_NT)), blk(TRSM_NT);
    KScope ks(s, "trsm");
    if (lower && unit) hipLaunchKernelGGL((trsm_leaf_kernel<T, true, true>), grid, blk, 0, s, n, nrhs, Tm, B);
    else if (lower) hipLaunchKernelGGL((trsm_leaf_kernel<T, true, false>), grid, blk, 0, s, n, nrhs, Tm, B);
    else if (unit) hipLaunchKernelGGL((trsm_leaf_kernel<T, false, true>), grid, blk, 0, s, n, nrhs, Tm, B);
    else hipLaunchKernelGGL((trsm_leaf_kernel<T, false, false>), grid, blk, 0, s, n, nrhs, Tm, B);
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
