// Internal entry points shared by the HIP sources and the C-ABI layer.
#pragma once
#include "common.hpp"

namespace mg {

// Folded tiles T'(r,k) = D_r T(r,k): block row r of the lower triangle is a
// row-major TB x (r TB) matrix holding T'(r, 0..r-1); of the upper triangle
// a row-major TB x (n - (r+1) TB) matrix holding T'(r, r+1..nb-1).
__host__ __device__ inline long fold_off(bool lower, int r, int n, int TB) {
  if (lower) return (long)TB * TB * ((long)r * (r - 1) / 2);
  return (long)TB * ((long)r * n - (long)TB * r * (r + 1) / 2);
}
__host__ __device__ inline long fold_ld(bool lower, int r, int n, int TB) {
  return lower ? (long)r * TB : (long)n - (long)(r + 1) * TB;
}

// Tp := D_r T(r, k) for every off-diagonal tile of block row r, all block
// rows of one triangle in one launch (gemm.hip).  D: the inverted TB x TB
// diagonal blocks, row-major, one after another; Tp laid out by fold_off /
// fold_ld.  Tiles, slab depth and the order of the sums are those of one
// gemm() call per block row, so the result is bit-identical to it.
template <typename T>
void gemm_fold_rows(hipStream_t s, bool lower, int n, int TB, View<const T> Tm, const T *D, T *Tp);

// ----------------------------------------------------------------- LU ----
// Row-major working layout: element (i,j) of the n x n matrix at W[i*ld+j].
// Partial pivoting with libmat's rules: the pivot of column k is the first
// row (lowest index) holding max|a(i,k)| for i >= k (mat.h:1247-1262); the
// last column (k = n-1) is never pivoted (mat.h:4214); a pivot with
// |a(k,k)| < eps is swapped in but neither scales its column nor updates the
// rest of its leaf block (mat.h:4234-4235).  ipiv[k] (device, int32) gets the
// global row swapped with row k.  `leaf` is the width of the column blocks
// factored by the pivoting kernel; leaf = 24 reproduces the reference's
// block boundaries for the eps-skip rule (MAT_PLU_BLOCK_SIZE, mat.h:4196).
struct LuOpts {
  // pivoting leaf width (the panel kernel takes up to 64 columns; measured
  // at n = 16384, 32-column leaves factor faster: 122 ms against 127 ms)
  int leaf = 32;
  Mul mul = Mul::Native;  // precision of the trailing updates (float only)
  // float only: tournament pivoting for leaves of at most 32 columns and
  // 16384 rows (tslu.hip) instead of libmat's partial pivoting
  bool tournament = false;
};
template <typename T>
void getrf_rowmajor(hipStream_t s, int n, T *W, long ld, int *ipiv, T eps, const LuOpts &o);
// Same, also maintaining perm (device, length n): perm[i] = original row now
// at row i (libmat's Perm, mat.h:382-385).
template <typename T>
void getrf_rowmajor_perm(hipStream_t s, int n, T *W, long ld, int *ipiv, int *perm, T eps,
                         const LuOpts &o);

// Number of k < n-1 with ipiv[k] != k, and libmat's Perm (p[i] = original
// row now at row i) from the sequential swaps (host arrays).
int ipiv_to_perm(int n, const int *ipiv, size_t *perm);

// Full-pivoting LU PAQ = LU on a column-major n x n matrix in place
// (mat_lu, x86 branch mat.h:4502-4557).  rec[3k..3k+2] = {pivot row,
// pivot col, skipped}; device arrays.
template <typename T>
void getrf_full(hipStream_t s, int n, T *A, long lda, int *rec, T eps);

// -------------------------------------------------------------------- QR --
// Householder QR in place on the column-major m x n R (m >= n, ld ldr);
// Q (m x m, ld m), when not null, receives the orthogonal factor.
// Columns with ||x|| < eps get no reflector.
template <typename T>
void geqrf(hipStream_t s, int m, int n, T *R, long ldr, T *Q, T eps);

// -------------------------------------------------------------- Cholesky --
// In-place lower Cholesky of a column-major matrix (only the lower triangle
// is read or written).  info (device int) gets j+1 for the first column j
// whose pivot is <= 0, else stays 0.
template <typename T>
void potrf_lower(hipStream_t s, int n, T *A, long lda, int *info, Mul mul = Mul::Native);

// ------------------------------------------------------------ triangular --
// B := op(T)^-1 B for a left-side triangular T (n x n, strided view) and
// B (n x nrhs, strided view).  lower/upper refers to op(T) as addressed by
// the view; unit ignores the diagonal.
template <typename T>
void trsm_left(hipStream_t s, bool lower, bool unit, int n, int nrhs, View<const T> Tm,
               View<T> B, Mul mul = Mul::Native);
// X := L^-1 B for a unit lower triangle of at most 256 rows whose 32 x 32
// diagonal blocks' inverses are given (row-major, block i at linv + 1024 i,
// as the tournament leaves write them); L and B row-major.
void trsm_lower_unit_linv(hipStream_t s, int nt, int nrhs, View<const float> L, const float *linv,
                          View<float> B);

// ------------------------------------------------------------ utilities --
template <typename T>
void transpose(hipStream_t s, int rows, int cols, const T *src, long lds, T *dst, long ldd);
// Packed row-major LU -> column-major unit-lower L and upper U (full n x n).
template <typename T>
void split_lu_rowmajor(hipStream_t s, int n, const T *W, long ld, T *L, T *U);
// Packed column-major LU -> column-major L and U.
template <typename T>
void split_lu_colmajor(hipStream_t s, int n, const T *A, long ld, T *L, T *U);
// Copy the lower triangle of column-major A into L and zero L's upper part.
template <typename T>
void copy_lower(hipStream_t s, int n, const T *A, long lda, T *L, long ldl);
// Zero the strict upper triangle of a column-major matrix.
template <typename T>
void zero_upper(hipStream_t s, int n, T *A, long lda);
// dst[i + c*ldd] = src[perm[i] + c*lds] (gather rows through a device perm).
template <typename T>
void gather_rows(hipStream_t s, int n, int nrhs, const T *src, long lds, const int *perm,
                 T *dst, long ldd);
// The same with a conversion: dst[i + c*ldd] = (To)src[perm[i] + c*lds].
template <typename Ti, typename To>
void gather_rows_cvt(hipStream_t s, int n, int nrhs, const Ti *src, long lds, const int *perm,
                     To *dst, long ldd);
template <typename Ti, typename To>
void convert(hipStream_t s, long count, const Ti *src, To *dst);
template <typename Ti, typename To>
void convert2d(hipStream_t s, int rows, int cols, const Ti *src, long lds, To *dst, long ldd);
void fill_int(hipStream_t s, int *p, long count, int value);
void fill_u32(hipStream_t s, unsigned *p, long count, unsigned value);

}  // namespace mg
