// Blocked triangular solves for a few right-hand sides (the substitutions
// of mat_solve / mat_solve_spd, mat.h:4584-4803, and every refinement step
// of the mixed-precision solvers).
//
// MI355X design: the recursion used for wide right-hand sides degenerates
// for 1..64 columns (one thread per column in the leaves).  Here the
// triangle is cut into TB x TB diagonal blocks that are inverted once per
// factorization (one workgroup per block, in LDS).  A solve is then nb+1
// launches; launch k updates every remaining row block with the block
// column just solved (X_r -= T(r,k) X_k, all row blocks in parallel, T
// streamed once from HBM) and the workgroup of the next row block finishes
// it with the inverted diagonal block (X_r = D_r^-1 X_r).  Each solve reads
// the triangle once.
#include "ops.hpp"

namespace mg {

template <typename T>
struct TriTB;
template <>
struct TriTB<float> {
  static constexpr int TB = 128;
};
template <>
struct TriTB<double> {
  static constexpr int TB = 64;
};

// Stage rows [r0, r0+nr) x cols [c0, c0+nc) of a strided view into LDS
// (dst[i*ldd + j]), zero outside the range.  Loads are issued in groups of
// CH before their LDS stores, so the global latency is paid once per group.
template <typename T, int TB, int NT>
__device__ inline void stage_tile(View<const T> V, int r0, int nr, int c0, int nc, T *dst, int ldd) {
  constexpr int PER = TB * TB / NT;
  constexpr int CH = PER < 16 ? PER : 16;
  const bool rowmaj = V.cs == 1;
  for (int g = 0; g < PER; g += CH) {
    T v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int idx = threadIdx.x + (g + u) * NT;
      const int i = rowmaj ? idx / TB : idx % TB;
      const int j = rowmaj ? idx % TB : idx / TB;
      v[u] = (i < nr && j < nc) ? V(r0 + i, c0 + j) : T(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int idx = threadIdx.x + (g + u) * NT;
      const int i = rowmaj ? idx / TB : idx % TB;
      const int j = rowmaj ? idx % TB : idx / TB;
      dst[i * ldd + j] = v[u];
    }
  }
}

// Inverse of diagonal block b (rows/cols b*TB ..), written row-major to
// D + b*TB*TB; a short last block is padded with the identity.  Right-looking
// sweep with all threads: X = I; for each pivot j, scale row j (non-unit),
// then X(i,:) -= T(i,j) X(j,:) for the rows below (lower) / above (upper).
template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void tri_inv_kernel(int n, View<const T> Tm, T *D) {
  constexpr int TB = TriTB<T>::TB;
  constexpr int LD = TB + 1;
  __shared__ T S[TB * LD];  // the block
  __shared__ T X[TB * LD];  // its inverse
  const int b = blockIdx.x, o = b * TB;
  const int nr = min(TB, n - o);
  stage_tile<T, TB, 256>(Tm, o, nr, o, nr, S, LD);
  for (int idx = threadIdx.x; idx < TB * TB; idx += 256) {
    const int i = idx / TB, j = idx % TB;
    X[i * LD + j] = (i == j) ? T(1) : T(0);
    if (i == j && (i >= nr || UNIT)) S[i * LD + j] = T(1);
  }
  __syncthreads();
  for (int s = 0; s < TB; ++s) {
    const int j = LOWER ? s : TB - 1 - s;
    if (!UNIT) {
      const T inv = T(1) / S[j * LD + j];
      for (int c = threadIdx.x; c < TB; c += 256) X[j * LD + c] *= inv;
      __syncthreads();
    }
    // rows i on the far side of j; row j of X is non-zero only on columns
    // c <= j (lower) / c >= j (upper)
    const int nrows = LOWER ? TB - 1 - j : j;
    const int ncols = LOWER ? j + 1 : TB - j;
    for (int idx = threadIdx.x; idx < nrows * ncols; idx += 256) {
      const int i = LOWER ? j + 1 + idx / ncols : idx / ncols;
      const int c = LOWER ? idx % ncols : j + idx % ncols;
      X[i * LD + c] -= S[i * LD + j] * X[j * LD + c];
    }
    __syncthreads();
  }
  T *Db = D + (long)b * TB * TB;
  for (int idx = threadIdx.x; idx < TB * TB; idx += 256) Db[idx] = X[(idx / TB) * LD + idx % TB];
}

// One block step of the forward (LOWER) or backward solve.
//   k: the block column solved by the previous launch (-1 / nb for the first)
//   workgroup w handles row block r = k+1+w (lower) or k-1-w (upper)
// The off-diagonal block and, for the finishing workgroup, the inverted
// diagonal block are staged through LDS with coalesced loads.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void tri_step_kernel(int n, int nrhs, int k, View<const T> Tm,
                                                       const T *D, T *X, long ldx) {
  constexpr int TB = TriTB<T>::TB;
  constexpr int NC = 16;  // right-hand sides per pass
  constexpr int RPT = TB * NC / 256;
  constexpr int LDT = TB + 1;
  __shared__ T Ts[TB * LDT];  // Ts[i*LDT + j]
  __shared__ T Xk[TB * NC];   // solved block column, [j][c]
  __shared__ T Y[TB * NC];    // this row block, [i][c]
  const int nb = (n + TB - 1) / TB;
  const int w = blockIdx.x;
  const int r = LOWER ? k + 1 + w : k - 1 - w;
  const bool finish = LOWER ? (r == k + 1) : (r == k - 1);
  const bool has_k = LOWER ? (k >= 0) : (k < nb);
  const int ro = r * TB, ko = k * TB;
  const int rrows = min(TB, n - ro);
  const int krows = has_k ? min(TB, n - ko) : 0;
  const int t = threadIdx.x;
  const int i = t % TB;           // row inside the block
  const int c0 = t / TB;          // first rhs column of this thread
  constexpr int CSTEP = 256 / TB;  // columns interleave
  if (has_k) stage_tile<T, TB, 256>(Tm, ro, rrows, ko, krows, Ts, LDT);
  for (int cb = 0; cb < nrhs; cb += NC) {
    const int nc = min(NC, nrhs - cb);
    if (has_k)
      for (int idx = t; idx < TB * NC; idx += 256) {
        const int j = idx % TB, c = idx / TB;
        Xk[j * NC + c] = (j < krows && c < nc) ? X[(ko + j) + (long)(cb + c) * ldx] : T(0);
      }
    __syncthreads();
    T acc[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int c = c0 + q * CSTEP;
      acc[q] = (i < rrows && c < nc) ? X[(ro + i) + (long)(cb + c) * ldx] : T(0);
    }
    if (has_k) {
      for (int j = 0; j < krows; ++j) {
        const T a = Ts[i * LDT + j];
#pragma unroll
        for (int q = 0; q < RPT; ++q) acc[q] -= a * Xk[j * NC + c0 + q * CSTEP];
      }
    }
    if (finish) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) Y[i * NC + c0 + q * CSTEP] = acc[q];
      __syncthreads();
      stage_tile<T, TB, 256>(View<const T>{D + (long)r * TB * TB, TB, 1}, 0, TB, 0, TB, Ts, LDT);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < RPT; ++q) acc[q] = T(0);
      for (int j = 0; j < TB; ++j) {
        const T a = Ts[i * LDT + j];
#pragma unroll
        for (int q = 0; q < RPT; ++q) acc[q] += a * Y[j * NC + c0 + q * CSTEP];
      }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int c = c0 + q * CSTEP;
      if (i < rrows && c < nc) X[(ro + i) + (long)(cb + c) * ldx] = acc[q];
    }
    __syncthreads();
    if (finish && has_k && cb + NC < nrhs) {
      // Ts now holds D_r^-1; restore T(r, k) for the next column pass
      stage_tile<T, TB, 256>(Tm, ro, rrows, ko, krows, Ts, LDT);
    }
  }
}

template <typename T>
int tri_block_size() {
  return TriTB<T>::TB;
}

// D must hold ceil(n/TB) * TB * TB elements.
template <typename T>
void tri_invert_blocks(hipStream_t s, bool lower, bool unit, int n, View<const T> Tm, T *D) {
  constexpr int TB = TriTB<T>::TB;
  const int nb = (n + TB - 1) / TB;
  if (nb == 0) return;
  if (lower && unit) hipLaunchKernelGGL((tri_inv_kernel<T, true, true>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else if (lower) hipLaunchKernelGGL((tri_inv_kernel<T, true, false>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else if (unit) hipLaunchKernelGGL((tri_inv_kernel<T, false, true>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else hipLaunchKernelGGL((tri_inv_kernel<T, false, false>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  MG_HIP(hipGetLastError());
}

// X := op(T)^-1 X, X column-major n x nrhs; D from tri_invert_blocks.
template <typename T>
void tri_solve_blocked(hipStream_t s, bool lower, int n, int nrhs, View<const T> Tm, const T *D,
                       T *X, long ldx) {
  constexpr int TB = TriTB<T>::TB;
  const int nb = (n + TB - 1) / TB;
  if (nb == 0 || nrhs <= 0) return;
  if (lower) {
    for (int k = -1; k < nb - 1; ++k)
      hipLaunchKernelGGL((tri_step_kernel<T, true>), dim3(nb - 1 - k), dim3(256), 0, s, n, nrhs, k, Tm,
                         D, X, ldx);
  } else {
    for (int k = nb; k > 0; --k)
      hipLaunchKernelGGL((tri_step_kernel<T, false>), dim3(k), dim3(256), 0, s, n, nrhs, k, Tm, D, X,
                         ldx);
  }
  MG_HIP(hipGetLastError());
}

#define MG_TRI_INST(T)                                                                           \
  template int tri_block_size<T>();                                                              \
  template void tri_invert_blocks<T>(hipStream_t, bool, bool, int, View<const T>, T *);          \
  template void tri_solve_blocked<T>(hipStream_t, bool, int, int, View<const T>, const T *, T *, \
                                     long);
MG_TRI_INST(float)
MG_TRI_INST(double)

}  // namespace mg
