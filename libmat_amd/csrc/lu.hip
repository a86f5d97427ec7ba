// LU drivers and layout utilities.
//
// getrf_rowmajor: recursive right-looking partial-pivot LU of the row-major
// working matrix (mat_plu's mat__plu_blocked_, mat.h:4201-4288, restated as
// a recursion so the trailing updates become a few large MFMA GEMMs
// instead of n/24 rank-24 updates).  The leaves are the pivoting kernel of
// panel.hip; every leaf applies its row interchanges to the whole row at
// once, so no deferred swap passes are needed.  The matrix is kept
// row-major on the device because a row interchange is then a contiguous
// (coalesced) move and each panel row is one contiguous run.
//
// getrf_full: full-pivoting LU PAQ = LU (mat_lu, x86 branch
// mat.h:4502-4557), one pivot search per step over the trailing
// submatrix, reproducing the reference's row-major scan order.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "ops.hpp"

namespace mg {

template <typename T>
void lu_leaf(hipStream_t s, int n, T *W, long ld, int c0, int w, int *ipiv, int *perm, int *pairs,
             T eps, int cl, int ch, bool tournament, T *linv, int prev_c0 = -1);
int tslu_max_rows();  // tallest panel of the tournament leaf (tslu.hip)
template <typename T>
void lu_rowswap(hipStream_t s, T *W, long ld, const int *pairs, int *perm, int k0, int w, int cl,
                int ch);
template <typename T>
void lu_rowswap_block(hipStream_t s, T *W, long ld, const int *pairs, int b0, int b1, int leaf,
                      int cl, int ch);
constexpr int kPairStride = 257;  // ints per leaf in the row-pair buffer (panel.hip)
int lu_leaf_max(int m, int elem_size);  // widest leaf the panel kernel takes for m rows

template <typename T>
static void lu_gemm_sub(hipStream_t s, long M, long N, long K, View<const T> A, View<const T> B,
                        View<T> C, Mul mul, int grid_cap = 0) {
  if constexpr (std::is_same<T, float>::value) {
    // small products: native fp32 MFMA is as fast as packing for the split
    // passes (and exact to fp32); the split pays off on large updates only
    if ((mul == Mul::Bf16x1 || mul == Mul::Bf16x3 || mul == Mul::Bf16x6) &&
        (double)M * N * K >= kLowpMinWork) {
      gemm_bf16x(s, M, N, K, -1.0f, A, B, C, (int)mul, false, grid_cap);
      return;
    }
  }
  gemm<T>(s, M, N, K, T(-1), A, B, T(1), C, false);
}

// Workgroups of the side stream's split-BF16 trailing updates per compute
// unit (MATGPU_SIDE_WG_PER_CU, a fraction allowed, 0 = uncapped): capped, so
// that the kernels of the critical path (panels, small solves and products)
// find free slots instead of waiting for GEMM tiles to finish.  A workgroup
// of the split GEMM takes 48 KB of LDS and 128 VGPRs per lane on each SIMD;
// two of them leave a unit 64 KB of LDS and 256 VGPRs per SIMD, which every
// critical-path kernel of the LU fits in (the tournament leaf 49 KB and 164
// VGPRs, the block-inverse solve 62 KB); a third leaves too little LDS for
// the block-inverse solve.
static int side_grid_cap() {
  static const double per_cu = [] {
    const char *e = std::getenv("MATGPU_SIDE_WG_PER_CU");
    return e ? std::max(0.0, std::atof(e)) : 2.0;
  }();
  return (int)(per_cu * num_cus());
}

// X := L^-1 X for the unit lower triangle of rows [r0, r0 + nt): from the
// tournament leaves' diagonal-block inverses when there are (linv, block
// solves in one launch), else by substitution.
template <typename T>
static void trsm_unit_lower(hipStream_t s, int nt, int nrhs, View<const T> L, View<T> X, T *linv,
                            int r0, Mul mul) {
  if constexpr (std::is_same<T, float>::value) {
    if (linv && nt <= 256) {
      trsm_lower_unit_linv(s, nt, nrhs, L, linv + (long)r0 * 32, X);
      return;
    }
  }
  trsm_left<T>(s, true, true, nt, nrhs, L, X, mul);
}

// Factor columns [c0, c0+w) (rows c0..n-1, all earlier updates applied)
// with leaves of at most `leaf` columns (every leaf starts at c0 + a
// multiple of `leaf`); the leaves' row interchanges reach columns [cl, ch).
// linv: where the tournament leaves put their diagonal blocks' inverses
// (32-column leaves), or null.
template <typename T>
static void getrf_rec(hipStream_t s, int n, T *W, long ld, int c0, int w, int *ipiv, int *perm,
                      int *pairs, T eps, const LuOpts &o, int leaf, int cl, int ch, T *linv) {
  if (w <= leaf) {
    lu_leaf<T>(s, n, W, ld, c0, w, ipiv, perm, pairs, eps, cl, ch, o.tournament, linv);
    return;
  }
  int w1 = ((w / 2 + leaf - 1) / leaf) * leaf;
  if (w1 >= w) w1 = w - leaf;
  const int w2 = w - w1;
  if (linv && leaf == 32 && w1 == leaf) {
    // a node of two 32-column tournament leaves: the second one also solves
    // U12 = L11^-1 A12 and updates its own rows (A22 -= L21 U12) in its launch
    lu_leaf<T>(s, n, W, ld, c0, w1, ipiv, perm, pairs, eps, cl, ch, o.tournament, linv);
    lu_leaf<T>(s, n, W, ld, c0 + w1, w2, ipiv, perm, pairs, eps, cl, ch, o.tournament, linv, c0);
    return;
  }
  getrf_rec<T>(s, n, W, ld, c0, w1, ipiv, perm, pairs, eps, o, leaf, cl, ch, linv);
  // U12 = L11^-1 A12 (unit lower)
  View<const T> L11{W + (long)c0 * ld + c0, ld, 1};
  View<T> A12{W + (long)c0 * ld + c0 + w1, ld, 1};
  trsm_unit_lower<T>(s, w1, w2, L11, A12, linv, c0, o.mul);
  // A22 -= L21 U12
  const int m2 = n - c0 - w1;
  View<const T> L21{W + (long)(c0 + w1) * ld + c0, ld, 1};
  View<T> A22{W + (long)(c0 + w1) * ld + c0 + w1, ld, 1};
  lu_gemm_sub<T>(s, m2, w2, w1, L21, cview(A12), A22, o.mul);
  getrf_rec<T>(s, n, W, ld, c0 + w1, w2, ipiv, perm, pairs, eps, o, leaf, cl, ch, linv);
}

// Row interchanges of the factored block [b0, b1) (its leaves in order)
// applied to columns [cl, ch).
template <typename T>
static void block_swaps(hipStream_t s, T *W, long ld, const int *pairs, int b0, int b1, int leaf,
                        int cl, int ch) {
  if (ch <= cl) return;
  lu_rowswap_block<T>(s, W, ld, pairs, b0, b1, leaf, cl, ch);
}

// U12 = L11^-1 A12 and A22 -= L21 U12 for the factored block [b0, b1) and
// the columns [c0, c1).
template <typename T>
static void block_update(hipStream_t s, int n, T *W, long ld, int b0, int b1, int c0, int c1,
                         const LuOpts &o, T *linv, int grid_cap = 0) {
  if (c1 <= c0) return;
  const int nb = b1 - b0;
  View<const T> L11{W + (long)b0 * ld + b0, ld, 1};
  View<T> A12{W + (long)b0 * ld + c0, ld, 1};
  trsm_unit_lower<T>(s, nb, c1 - c0, L11, A12, linv, b0, o.mul);
  View<const T> L21{W + (long)b1 * ld + b0, ld, 1};
  View<T> A22{W + (long)b1 * ld + c0, ld, 1};
  lu_gemm_sub<T>(s, n - b1, c1 - c0, nb, L21, cview(A12), A22, o.mul, grid_cap);
}

// Right-looking blocked LU with look-ahead over NB-wide column blocks.
// The caller's stream (the critical path) factors block k+1 right after
// applying block k to block k+1's columns; a second stream meanwhile
// applies block k to everything right of block k+1 (interchanges, TRSM,
// the large GEMM) and then brings block k's interchanges to the columns
// left of block k.  Within a block the recursion's leaves swap only the
// block's own columns.  The two streams never touch the same columns at
// the same time:
//   caller:  panel(k+1) on [b1, c2)   after the side stream has applied
//            block k-1 to those columns (the first part of its update)
//   side:    block k on [c2, c3), then on [c3, n), then the interchanges on
//            [0, b0)   after panel(k) is done
// and interchanges of different blocks reach any column in pivot order.
template <typename T>
static void getrf_lookahead(hipStream_t s, int n, T *W, long ld, int *ipiv, int *perm, int *pairs,
                            T eps, const LuOpts &o, int NB, T *linv) {
  // the side stream and the event pool are per device: one enqueue at a time
  std::lock_guard<std::mutex> lk(lookahead_mutex());
  hipStream_t u = side_stream();
  const int nblk = (n + NB - 1) / NB;
  std::vector<hipEvent_t> evP(nblk), evU(nblk);
  for (int k = 0; k < nblk; ++k) {
    evP[k] = pooled_event(2 * k);
    evU[k] = pooled_event(2 * k + 1);
  }
  // the side stream starts after everything queued on the caller's stream
  hipEvent_t start = pooled_event(2 * nblk);
  MG_HIP(hipEventRecord(start, s));
  MG_HIP(hipStreamWaitEvent(u, start, 0));
  // leaf width of a block: the panel kernel's limit for the block's height
  auto leaf_of = [&](int b0) { return std::min(o.leaf, lu_leaf_max(n - b0, (int)sizeof(T))); };
  getrf_rec<T>(s, n, W, ld, 0, std::min(NB, n), ipiv, perm, pairs, eps, o, leaf_of(0), 0,
               std::min(NB, n), linv);
  for (int k = 0; k < nblk; ++k) {
    const int b0 = k * NB, b1 = std::min(n, b0 + NB);
    const int c2 = std::min(n, b1 + NB);
    MG_HIP(hipEventRecord(evP[k], s));
    if (b1 < n) {
      // caller: block k+1's columns, then its panel
      if (k >= 1) MG_HIP(hipStreamWaitEvent(s, evU[k - 1], 0));
      block_swaps<T>(s, W, ld, pairs, b0, b1, leaf_of(b0), b1, c2);
      block_update<T>(s, n, W, ld, b0, b1, b1, c2, o, linv);
      getrf_rec<T>(s, n, W, ld, b1, c2 - b1, ipiv, perm, pairs, eps, o, leaf_of(b1), b1, c2,
                   linv);
    }
    // side: the rest of the trailing matrix, block k+2's columns first (the
    // only part the caller's next step waits for), then the columns on the left
    MG_HIP(hipStreamWaitEvent(u, evP[k], 0));
    const int c3 = std::min(n, c2 + NB);
    block_swaps<T>(u, W, ld, pairs, b0, b1, leaf_of(b0), c2, n);
    block_update<T>(u, n, W, ld, b0, b1, c2, c3, o, linv, side_grid_cap());
    MG_HIP(hipEventRecord(evU[k], u));
    block_update<T>(u, n, W, ld, b0, b1, c3, n, o, linv, side_grid_cap());
    block_swaps<T>(u, W, ld, pairs, b0, b1, leaf_of(b0), 0, b0);
  }
  hipEvent_t done = pooled_event(2 * nblk + 1);
  MG_HIP(hipEventRecord(done, u));
  MG_HIP(hipStreamWaitEvent(s, done, 0));
}

__global__ void iota_kernel(int *p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = i;
}

// ipiv and perm are device int arrays of length n; on return perm[i] is
// the original row now at row i (libmat's Perm, mat.h:382-385).
template <typename T>
void getrf_rowmajor_perm(hipStream_t s, int n, T *W, long ld, int *ipiv, int *perm, T eps,
                         const LuOpts &o) {
  MG_CHECK(o.leaf >= 1 && o.leaf <= 64, "getrf: leaf width must be in [1, 64]");
  hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, ipiv, n);
  hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, perm, n);
  MG_HIP(hipGetLastError());
  // every leaf's interchanges as row pairs, written by its panel kernel
  int *pairs = (int *)workspace(s).get("getrf.pairs", (size_t)n * kPairStride * sizeof(int));
  // the tournament leaves (fp32, 32 columns) also write the inverses of
  // their unit lower diagonal blocks, and the solves with the factored
  // blocks' L then run as block steps (trsm_lower_unit_linv)
  T *linv = nullptr;
  if (std::is_same<T, float>::value && o.tournament && o.leaf == 32 && n <= tslu_max_rows())
    linv = (T *)workspace(s).get("getrf.linv", (size_t)ceil_div(n, 32) * 32 * 32 * sizeof(T));
  const int NB = lookahead_block();
  if (NB > 0 && n >= 4 * NB && NB % o.leaf == 0)
    getrf_lookahead<T>(s, n, W, ld, ipiv, perm, pairs, eps, o, NB, linv);
  else
    getrf_rec<T>(s, n, W, ld, 0, n, ipiv, perm, pairs, eps, o,
                 std::min(o.leaf, lu_leaf_max(n, (int)sizeof(T))), 0, n, linv);
}

template <typename T>
void getrf_rowmajor(hipStream_t s, int n, T *W, long ld, int *ipiv, T eps, const LuOpts &o) {
  int *perm = (int *)workspace(s).get("getrf.perm", (size_t)n * sizeof(int));
  getrf_rowmajor_perm<T>(s, n, W, ld, ipiv, perm, eps, o);
}

int ipiv_to_perm(int n, const int *ipiv, size_t *perm) {
  for (int i = 0; i < n; ++i) perm[i] = (size_t)i;
  int swaps = 0;
  for (int k = 0; k + 1 < n; ++k) {
    const int p = ipiv[k];
    if (p != k) {
      size_t t = perm[k];
      perm[k] = perm[p];
      perm[p] = t;
      ++swaps;
    }
  }
  return swaps;
}

// ------------------------------------------------------ full pivoting ----
// Per-workgroup partial maxima over the trailing submatrix; key order:
// larger |v| first, then smaller row, then smaller column (the reference's
// i-outer / j-inner scan with strict '>' keeps the first maximum).
template <typename T>
__device__ inline void fp_better(T &bv, int &bi, int &bj, T v, int i, int j) {
  if (v > bv || (v == bv && (i < bi || (i == bi && j < bj)))) {
    bv = v;
    bi = i;
    bj = j;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void fp_amax_kernel(int n, int k, const T *A, long lda, T *pv,
                                                      int *pi, int *pj) {
  const int m = n - k;
  const long total = (long)m * m;
  T bv = T(-1);
  int bi = 0x7fffffff, bj = 0x7fffffff;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int i = k + (int)(idx % m), j = k + (int)(idx / m);
    fp_better(bv, bi, bj, (T)fabs(A[i + (long)j * lda]), i, j);
  }
  __shared__ T sv[256];
  __shared__ int si[256], sj[256];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  sj[threadIdx.x] = bj;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      T v = sv[threadIdx.x];
      int i = si[threadIdx.x], j = sj[threadIdx.x];
      fp_better(v, i, j, sv[threadIdx.x + off], si[threadIdx.x + off], sj[threadIdx.x + off]);
      sv[threadIdx.x] = v;
      si[threadIdx.x] = i;
      sj[threadIdx.x] = j;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
    pj[blockIdx.x] = sj[0];
  }
}

template <typename T>
__global__ void fp_select_kernel(int nparts, int k, const T *pv, const int *pi, const int *pj,
                                 const T *A, long lda, T eps, int *rec) {
  if (threadIdx.x != 0) return;
  // the reference starts from |A(k,k)| and only replaces on strictly greater
  T bv = (T)fabs(A[k + (long)k * lda]);
  int bi = k, bj = k;
  for (int x = 0; x < nparts; ++x)
    if (pv[x] > bv) {
      bv = pv[x];
      bi = pi[x];
      bj = pj[x];
    } else if (pv[x] == bv && (pi[x] < bi || (pi[x] == bi && pj[x] < bj))) {
      bi = pi[x];
      bj = pj[x];
    }
  const bool skip = bv < eps;  // also false for NaN, as in the reference
  rec[3 * k] = bi;
  rec[3 * k + 1] = bj;
  rec[3 * k + 2] = skip ? 1 : 0;
}

template <typename T>
__global__ void fp_swap_rows_kernel(int n, int k, T *A, long lda, const int *rec) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (rec[3 * k + 2] || j >= n) return;
  const int pr = rec[3 * k];
  if (pr == k) return;
  T t = A[k + (long)j * lda];
  A[k + (long)j * lda] = A[pr + (long)j * lda];
  A[pr + (long)j * lda] = t;
}

template <typename T>
__global__ void fp_swap_cols_kernel(int n, int k, T *A, long lda, const int *rec) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (rec[3 * k + 2] || i >= n) return;
  const int pc = rec[3 * k + 1];
  if (pc == k) return;
  T t = A[i + (long)k * lda];
  A[i + (long)k * lda] = A[i + (long)pc * lda];
  A[i + (long)pc * lda] = t;
}

template <typename T>
__global__ void fp_update_kernel(int n, int k, T *A, long lda, const int *rec) {
  if (rec[3 * k + 2]) return;
  const int m = n - k - 1;
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * m) return;
  const int i = k + 1 + (int)(idx % m), j = k + 1 + (int)(idx / m);
  const T inv = T(1) / A[k + (long)k * lda];
  const T l = A[i + (long)k * lda] * inv;
  A[i + (long)j * lda] -= l * A[k + (long)j * lda];
}

template <typename T>
__global__ void fp_scale_kernel(int n, int k, T *A, long lda, const int *rec) {
  if (rec[3 * k + 2]) return;
  const int i = k + 1 + blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const T inv = T(1) / A[k + (long)k * lda];
  A[i + (long)k * lda] = A[i + (long)k * lda] * inv;
}

template <typename T>
void getrf_full(hipStream_t s, int n, T *A, long lda, int *rec, T eps) {
  Workspace &ws = workspace(s);
  const int maxparts = 1024;
  T *pv = (T *)ws.get("fp.pv", maxparts * sizeof(T));
  int *pi = (int *)ws.get("fp.pi", maxparts * sizeof(int));
  int *pj = (int *)ws.get("fp.pj", maxparts * sizeof(int));
  for (int k = 0; k + 1 < n; ++k) {
    const long m = n - k;
    const int parts = (int)std::min<long>(maxparts, (m * m + 255) / 256);
    hipLaunchKernelGGL((fp_amax_kernel<T>), dim3(parts), dim3(256), 0, s, n, k, A, lda, pv, pi, pj);
    hipLaunchKernelGGL((fp_select_kernel<T>), dim3(1), dim3(64), 0, s, parts, k, pv, pi, pj, A, lda,
                       eps, rec);
    hipLaunchKernelGGL((fp_swap_rows_kernel<T>), dim3(ceil_div(n, 256)), dim3(256), 0, s, n, k, A,
                       lda, rec);
    hipLaunchKernelGGL((fp_swap_cols_kernel<T>), dim3(ceil_div(n, 256)), dim3(256), 0, s, n, k, A,
                       lda, rec);
    const long mm = (m - 1) * (m - 1);
    if (mm > 0)
      hipLaunchKernelGGL((fp_update_kernel<T>), dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, s,
                         n, k, A, lda, rec);
    hipLaunchKernelGGL((fp_scale_kernel<T>), dim3(ceil_div(m, 256)), dim3(256), 0, s, n, k, A, lda,
                       rec);
    MG_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------------ utilities --
template <typename T>
__global__ void transpose_kernel(int rows, int cols, const T *src, long lds, T *dst, long ldd) {
  // src: rows x cols column-major (ld lds); dst: cols x rows column-major
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    const int i = bx + tx, j = by + yy;
    if (i < rows && j < cols) tile[yy][tx] = src[i + (long)j * lds];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int j = by + tx, i = bx + yy;
    if (i < rows && j < cols) dst[j + (long)i * ldd] = tile[tx][yy];
  }
}

template <typename T>
void transpose(hipStream_t s, int rows, int cols, const T *src, long lds, T *dst, long ldd) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL((transpose_kernel<T>), dim3(ceil_div(rows, 32), ceil_div(cols, 32)), dim3(256),
                     0, s, rows, cols, src, lds, dst, ldd);
  MG_HIP(hipGetLastError());
}

template <typename T>
__global__ void split_rm_kernel(int n, const T *W, long ld, T *L, T *U) {
  __shared__ T tile[32][33];
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;  // tile rows i, cols j
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int i = bi + yy, j = bj + tx;
    if (i < n && j < n) tile[yy][tx] = W[(long)i * ld + j];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int j = bj + yy, i = bi + tx;
    if (i < n && j < n) {
      const T v = tile[tx][yy];
      const long o = i + (long)j * n;
      L[o] = (i > j) ? v : (i == j ? T(1) : T(0));
      U[o] = (i <= j) ? v : T(0);
    }
  }
}

template <typename T>
void split_lu_rowmajor(hipStream_t s, int n, const T *W, long ld, T *L, T *U) {
  hipLaunchKernelGGL((split_rm_kernel<T>), dim3(ceil_div(n, 32), ceil_div(n, 32)), dim3(256), 0, s,
                     n, W, ld, L, U);
  MG_HIP(hipGetLastError());
}

template <typename T>
__global__ void split_cm_kernel(int n, const T *A, long ld, T *L, T *U) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const T v = A[i + (long)j * ld];
  L[idx] = (i > j) ? v : (i == j ? T(1) : T(0));
  U[idx] = (i <= j) ? v : T(0);
}

template <typename T>
void split_lu_colmajor(hipStream_t s, int n, const T *A, long ld, T *L, T *U) {
  hipLaunchKernelGGL((split_cm_kernel<T>), dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0, s, n,
                     A, ld, L, U);
  MG_HIP(hipGetLastError());
}

// One workgroup per column j: L(:, j) = A(:, j) on and below the diagonal,
// zero above it; with VEC, 16-byte accesses (both columns 16-byte aligned,
// n a multiple of the vector width), and the part above the diagonal is
// written without being read.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void copy_lower_kernel(int n, const T *A, long lda, T *L,
                                                         long ldl) {
  const int j = blockIdx.x;
  const T *a = A + (long)j * lda;
  T *l = L + (long)j * ldl;
  if constexpr (VEC) {
    constexpr int V = 16 / (int)sizeof(T);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    for (int i0 = threadIdx.x * V; i0 < n; i0 += 256 * V) {
      Vec v;
      if (i0 + V - 1 < j) {
        v = Vec{};
      } else {
        v = *reinterpret_cast<const Vec *>(a + i0);
        T *e = reinterpret_cast<T *>(&v);
#pragma unroll
        for (int k = 0; k < V; ++k)
          if (i0 + k < j) e[k] = T(0);
      }
      *reinterpret_cast<Vec *>(l + i0) = v;
    }
  } else {
    for (int i = threadIdx.x; i < n; i += 256) l[i] = i >= j ? a[i] : T(0);
  }
}

template <typename T>
void copy_lower(hipStream_t s, int n, const T *A, long lda, T *L, long ldl) {
  if (n <= 0) return;
  constexpr int V = 16 / (int)sizeof(T);
  const bool vec = n % V == 0 && lda % V == 0 && ldl % V == 0 && (uintptr_t)A % 16 == 0 &&
                   (uintptr_t)L % 16 == 0;
  if (vec)
    hipLaunchKernelGGL((copy_lower_kernel<T, true>), dim3(n), dim3(256), 0, s, n, A, lda, L, ldl);
  else
    hipLaunchKernelGGL((copy_lower_kernel<T, false>), dim3(n), dim3(256), 0, s, n, A, lda, L, ldl);
  MG_HIP(hipGetLastError());
}

template <typename T>
__global__ void zero_upper_kernel(int n, T *A, long lda) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i < j) A[i + (long)j * lda] = T(0);
}

template <typename T>
void zero_upper(hipStream_t s, int n, T *A, long lda) {
  hipLaunchKernelGGL((zero_upper_kernel<T>), dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0, s, n,
                     A, lda);
  MG_HIP(hipGetLastError());
}

template <typename Ti, typename To>
__global__ void gather_rows_kernel(int n, int nrhs, const Ti *src, long lds, const int *perm,
                                   To *dst, long ldd) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * nrhs) return;
  const int i = (int)(idx % n), c = (int)(idx / n);
  dst[i + (long)c * ldd] = (To)src[perm[i] + (long)c * lds];
}

template <typename Ti, typename To>
void gather_rows_cvt(hipStream_t s, int n, int nrhs, const Ti *src, long lds, const int *perm,
                     To *dst, long ldd) {
  if (n <= 0 || nrhs <= 0) return;
  hipLaunchKernelGGL((gather_rows_kernel<Ti, To>), dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256),
                     0, s, n, nrhs, src, lds, perm, dst, ldd);
  MG_HIP(hipGetLastError());
}

template <typename T>
void gather_rows(hipStream_t s, int n, int nrhs, const T *src, long lds, const int *perm, T *dst,
                 long ldd) {
  gather_rows_cvt<T, T>(s, n, nrhs, src, lds, perm, dst, ldd);
}

template <typename Ti, typename To>
__global__ void convert_kernel(long count, const Ti *src, To *dst) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long)gridDim.x * 256)
    dst[i] = (To)src[i];
}

template <typename Ti, typename To>
void convert(hipStream_t s, long count, const Ti *src, To *dst) {
  if (count <= 0) return;
  const unsigned grid = (unsigned)std::min<long>((count + 255) / 256, 65536);
  hipLaunchKernelGGL((convert_kernel<Ti, To>), dim3(grid), dim3(256), 0, s, count, src, dst);
  MG_HIP(hipGetLastError());
}

template <typename Ti, typename To>
__global__ void convert2d_kernel(int rows, int cols, const Ti *src, long lds, To *dst, long ldd) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)rows * cols) return;
  const int i = (int)(idx % rows), j = (int)(idx / rows);
  dst[i + (long)j * ldd] = (To)src[i + (long)j * lds];
}

template <typename Ti, typename To>
void convert2d(hipStream_t s, int rows, int cols, const Ti *src, long lds, To *dst, long ldd) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL((convert2d_kernel<Ti, To>), dim3(ceil_div((size_t)rows * cols, 256)), dim3(256),
                     0, s, rows, cols, src, lds, dst, ldd);
  MG_HIP(hipGetLastError());
}

__global__ void fill_int_kernel(int *p, long count, int v) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}
__global__ void fill_u32_kernel(unsigned *p, long count, unsigned v) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}
void fill_int(hipStream_t s, int *p, long count, int value) {
  hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, value);
  MG_HIP(hipGetLastError());
}
void fill_u32(hipStream_t s, unsigned *p, long count, unsigned value) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, value);
  MG_HIP(hipGetLastError());
}

#define MG_INST(T)                                                                              \
  template void getrf_rowmajor<T>(hipStream_t, int, T *, long, int *, T, const LuOpts &);       \
  template void getrf_rowmajor_perm<T>(hipStream_t, int, T *, long, int *, int *, T,             \
                                       const LuOpts &);                                          \
  template void getrf_full<T>(hipStream_t, int, T *, long, int *, T);                           \
  template void transpose<T>(hipStream_t, int, int, const T *, long, T *, long);                \
  template void split_lu_rowmajor<T>(hipStream_t, int, const T *, long, T *, T *);              \
  template void split_lu_colmajor<T>(hipStream_t, int, const T *, long, T *, T *);              \
  template void copy_lower<T>(hipStream_t, int, const T *, long, T *, long);                    \
  template void zero_upper<T>(hipStream_t, int, T *, long);                                     \
  template void gather_rows<T>(hipStream_t, int, int, const T *, long, const int *, T *, long);
MG_INST(float)
MG_INST(double)
template void convert<double, float>(hipStream_t, long, const double *, float *);
template void convert<float, double>(hipStream_t, long, const float *, double *);
template void gather_rows_cvt<double, float>(hipStream_t, int, int, const double *, long,
                                              const int *, float *, long);
template void convert2d<double, float>(hipStream_t, int, int, const double *, long, float *, long);
template void convert2d<float, double>(hipStream_t, int, int, const float *, long, double *, long);
template void convert2d<double, double>(hipStream_t, int, int, const double *, long, double *, long);
template void convert2d<float, float>(hipStream_t, int, int, const float *, long, float *, long);

}  // namespace mg
