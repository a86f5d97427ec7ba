/*
 * mat_gpu.h - C ABI of libmat_amd, the MI355X (gfx950) implementation of
 * libmat's dense GEMM + factor-and-solve path.
 *
 * Two shared libraries carry the same symbol names, one per element type,
 * exactly as libmat selects mat_elem_t at compile time (mat.h:165-181):
 *     libmat_amd/libmatgpu_f32.so   (float,  eps 1e-6f)
 *     libmat_amd/libmatgpu_f64.so   (double, eps 1e-9; MAT_DOUBLE_PRECISION)
 *
 * Part 1 is the drop-in: the reference's own prototypes, same structs,
 * same column-major storage (mat.h:387-393), same pivot output and return
 * codes.  A libmat user includes mat.h *without* MAT_IMPLEMENTATION and
 * links one of the libraries above.  These entry points take host `Mat`s,
 * run on the GPU and return synchronously, like the reference.  Contract
 * violations that the reference catches with MAT_ASSERT (mat.h:327-357)
 * abort with a message; there is no CPU fallback.
 *
 * Part 2 is the device-resident extension API (no counterpart in the
 * reference): plain device pointers, column-major, explicit leading
 * dimensions, a HIP stream passed as void* (NULL = the HIP null stream).
 * They return 0 on success, a negative code on error (see
 * mat_gpu_last_error()), and -1 from the Cholesky routines when the matrix
 * is not positive definite, as mat_chol does.
 */
#ifndef MAT_GPU_H_
#define MAT_GPU_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types --
 * Identical to mat.h:165-181 and mat.h:364-385; skipped when mat.h has
 * already been included. */
#ifndef MAT_H_
#ifdef MAT_DOUBLE_PRECISION
typedef double mat_elem_t;
#else
typedef float mat_elem_t;
#endif
typedef struct {
  size_t rows;
  size_t cols;
  mat_elem_t *data; /* column-major: data[j * rows + i] */
} Mat;
typedef Mat Vec;
typedef struct {
  size_t *data; /* p[i] = original row now at row i */
  size_t size;
} Perm;
#endif

/* ================================================= Part 1: drop-in ====== */

/* out = a * b.                     replaces mat.h:564  (impl 1842-1850) */
void mat_mul(Mat *out, const Mat *a, const Mat *b);
/* new matrix (calloc'd, free with mat_free_mat) = a * b.
                                    replaces mat.h:567  (impl 1852-1862) */
Mat *mat_rmul(const Mat *a, const Mat *b);
/* C = alpha * A * B + beta * C.    replaces mat.h:612  (impl 2787-2796) */
void mat_gemm(Mat *C, mat_elem_t alpha, const Mat *A, const Mat *B, mat_elem_t beta);
/* C = alpha * A * A^T + beta * C on the uplo triangle ('L'/'l' lower, else
   upper; the other triangle is left as it is).
                                    replaces mat.h:617  (impl 5670-5677) */
void mat_syrk(Mat *C, const Mat *A, mat_elem_t alpha, mat_elem_t beta, char uplo);
/* C = alpha * A^T * A + beta * C on the uplo triangle.
                                    replaces mat.h:622  (impl 5727-5734) */
void mat_syrk_t(Mat *C, const Mat *A, mat_elem_t alpha, mat_elem_t beta, char uplo);
/* A = Q R by Householder reflections, A m x n with m >= n, Q m x m, R m x n
   (R keeps the reflectors' rounding residue below its diagonal, as libmat
   leaves it).                      replaces mat.h:744  (impl 3907-4063) */
void mat_qr(const Mat *A, Mat *Q, Mat *R);
/* R of A = Q R only.               replaces mat.h:749  (impl 4066-4192) */
void mat_qr_r(const Mat *A, Mat *R);
/* PAQ = LU, full pivoting; returns row swaps + column swaps.
                                    replaces mat.h:771  (impl 4476-4578) */
int mat_lu(const Mat *A, Mat *L, Mat *U, Perm *p, Perm *q);
/* PA = LU, partial pivoting; returns the number of row swaps.
                                    replaces mat.h:776  (impl 4436-4474) */
int mat_plu(const Mat *A, Mat *L, Mat *U, Perm *p);
/* A = L L^T from A's lower triangle; -1 if not positive definite.
                                    replaces mat.h:780  (impl 6062-6095) */
int mat_chol(const Mat *A, Mat *L);
/* A x = b via PLU.                 replaces mat.h:823  (impl 4807-4837) */
void mat_solve(Vec *x, const Mat *A, const Vec *b);
/* A x = b via Cholesky; -1 (x untouched) if A is not SPD.
                                    replaces mat.h:828  (impl 4839-4871) */
int mat_solve_spd(Vec *x, const Mat *A, const Vec *b);
/* triangular substitutions.        replace mat.h:832-838 (impl 4580-4803) */
void mat_solve_tril(Vec *x, const Mat *L, const Vec *b);
void mat_solve_tril_unit(Vec *x, const Mat *L, const Vec *b);
void mat_solve_triu(Vec *x, const Mat *U, const Vec *b);
void mat_solve_trilt(Vec *x, const Mat *L, const Vec *b);
/* determinant via PLU, inverse via PLU + solves.
                                    replace mat.h:734,789 (impl 6166-6249) */
mat_elem_t mat_det(const Mat *A);
void mat_inv(Mat *out, const Mat *A);

/* Host helpers a libmat caller uses around the path (mat.h:398-480), with
 * the reference's semantics (row-major initializer arrays, calloc'd data). */
Mat *mat_empty(size_t rows, size_t cols);
Mat *mat_mat(size_t rows, size_t cols);
Mat *mat_from(size_t rows, size_t cols, const mat_elem_t *values);
void mat_init(Mat *out, const mat_elem_t *values);
void mat_free_mat(Mat *m);
Mat *mat_zeros(size_t rows, size_t cols);
void mat_eye(Mat *out);
Mat *mat_reye(size_t dim);
Vec *mat_vec(size_t dim);
Vec *mat_vec_from(size_t dim, const mat_elem_t *values);
Perm *mat_perm(size_t n);
void mat_free_perm(Perm *p);
void mat_perm_identity(Perm *p);
Mat *mat_perm_mat(const Perm *p);
void mat_deep_copy(Mat *out, const Mat *src);
Mat *mat_rdeep_copy(const Mat *m);
mat_elem_t mat_at(const Mat *m, size_t row, size_t col);
void mat_set_at(Mat *m, size_t row, size_t col, mat_elem_t value);
bool mat_equals(const Mat *a, const Mat *b);                          /* mat.h:475 */
bool mat_equals_tol(const Mat *a, const Mat *b, mat_elem_t epsilon);  /* mat.h:478 */

/* ====================================== Part 2: device-resident API ====== */

/* Message of the last failed extension call on this thread ("" if none). */
const char *mat_gpu_last_error(void);
/* Element type of this library build: 4 (float) or 8 (double). */
int mat_gpu_elem_size(void);
/* Synchronizes the stream; returns non-zero (message in
 * mat_gpu_last_error()) if a multi-workgroup kernel queued on it since the
 * last check gave up a bounded wait for another workgroup.  Synchronous
 * entry points make this check themselves; asynchronous ones (getrs,
 * potrs, gesv, posv on a stream) are checked by the next synchronous call
 * on the same stream or by this one. */
int mat_gpu_stream_check(void *stream);
/* Opt-in kernel timing.  While enabled, the library records a HIP event
 * pair on the launch stream around each launch of these families and
 * files it with the launch's algorithmic work:
 *   "panel"   GEPP panel (bytes: the m x w panel read and written once)
 *   "rowswap" row interchanges outside the panel (work 0)
 *   "gemm"    native fp32/fp64 MFMA GEMM (flops 2mnk)
 *   "gemm_bf16x" split-bf16 GEMM (MFMA flops: passes * 2mnk)
 *   "gemm_fp8"   fp8 GEMM (flops 2mnk)
 *   "pack"    operand packing for the low-precision GEMMs (bytes moved)
 *   "trsm"    triangular-solve leaves and block inverses (work 0)
 *   "tri_chain" chained triangular solve (bytes: triangle + solutions)
 * Enabling clears earlier records (and synchronizes the device). */
int mat_gpu_timing_enable(int on);
/* Launch count, summed event time (ms) and summed work of one family;
 * synchronizes on the recorded events. */
int mat_gpu_timing_query(const char *family, int64_t *launches, double *total_ms, double *work);

/* C = alpha * op(A) * op(B) + beta * C; trans_* = 0 (N) or 1 (T).
 * MFMA fp32 / fp64.  beta == 0 never reads C.  Asynchronous on stream. */
int mat_gpu_gemm_f32(void *stream, int trans_a, int trans_b, size_t m, size_t n, size_t k,
                     float alpha, const float *a, size_t lda, const float *b, size_t ldb,
                     float beta, float *c, size_t ldc);
int mat_gpu_gemm_f64(void *stream, int trans_a, int trans_b, size_t m, size_t n, size_t k,
                     double alpha, const double *a, size_t lda, const double *b, size_t ldb,
                     double beta, double *c, size_t ldc);

/* C += alpha * A * B (all column-major fp32) on low-precision MFMA:
 * passes 1/3/6 = BF16 with 1/3/6 exact-split products, 8 = FP8 e4m3 with
 * one power-of-two scale per 128x128 tile; fp32 accumulation.  This is the
 * trailing-update multiply of the mixed-precision factorizations. */
int mat_gpu_gemm_lowp(void *stream, int passes, size_t m, size_t n, size_t k, float alpha,
                      const float *a, size_t lda, const float *b, size_t ldb, float *c, size_t ldc);

/* In-place partial-pivot LU of the n x n column-major a (packed: unit-lower
 * L below the diagonal, U on and above), libmat's pivot rules and eps-skip.
 * perm (device, n int32) gets libmat's Perm; *swaps (host) the swap count.
 * leaf: pivoting block width in [1, 64] (0: 32); 24 reproduces libmat's
 * block boundaries for near-zero pivots.  Synchronous. */
int mat_gpu_getrf_f32(void *stream, size_t n, float *a, size_t lda, int32_t *perm, int leaf,
                      int *swaps);
int mat_gpu_getrf_f64(void *stream, size_t n, double *a, size_t lda, int32_t *perm, int leaf,
                      int *swaps);
/* Same packed result with tournament pivoting (CALU) in the leaves of at
 * most 65536 rows: PA = LU with other pivots than libmat's partial
 * pivoting, and the workgroups meeting once per level of a reduction tree
 * per 32-column leaf instead of once per column.  No counterpart in libmat
 * (its LU is mat_plu, mat.h:4436-4474); the mixed solver's factorization
 * (pivoting = 2, the default). */
int mat_gpu_getrf_tournament_f32(void *stream, size_t n, float *a, size_t lda, int32_t *perm,
                                 int *swaps);
/* X = A^-1 B from mat_gpu_getrf's packed factor and perm; nrhs columns. */
int mat_gpu_getrs_f32(void *stream, size_t n, size_t nrhs, const float *lu, size_t ldlu,
                      const int32_t *perm, const float *b, size_t ldb, float *x, size_t ldx);
int mat_gpu_getrs_f64(void *stream, size_t n, size_t nrhs, const double *lu, size_t ldlu,
                      const int32_t *perm, const double *b, size_t ldb, double *x, size_t ldx);
/* In-place lower Cholesky (upper triangle zeroed); 0, or -1 if not SPD.
 * Synchronous. */
int mat_gpu_potrf_f32(void *stream, size_t n, float *a, size_t lda);
int mat_gpu_potrf_f64(void *stream, size_t n, double *a, size_t lda);
/* Householder QR in place on the column-major m x n a (m >= n): a becomes
 * mat_qr's R; q (m x m, ld m, device) receives Q unless null.  The device
 * form of mat_qr / mat_qr_r (mat.h:3907-4192).  Synchronous. */
int mat_gpu_geqrf_f32(void *stream, size_t m, size_t n, float *a, size_t lda, float *q);
int mat_gpu_geqrf_f64(void *stream, size_t m, size_t n, double *a, size_t lda, double *q);
/* X = (L L^T)^-1 B from mat_gpu_potrf's factor. */
int mat_gpu_potrs_f32(void *stream, size_t n, size_t nrhs, const float *l, size_t ldl,
                      const float *b, size_t ldb, float *x, size_t ldx);
int mat_gpu_potrs_f64(void *stream, size_t n, size_t nrhs, const double *l, size_t ldl,
                      const double *b, size_t ldb, double *x, size_t ldx);
/* Factor once + solve nrhs columns, A untouched (mat_solve / mat_solve_spd
 * generalised to many right-hand sides).  posv returns -1 if not SPD. */
int mat_gpu_gesv_f32(void *stream, size_t n, size_t nrhs, const float *a, size_t lda,
                     const float *b, size_t ldb, float *x, size_t ldx);
int mat_gpu_gesv_f64(void *stream, size_t n, size_t nrhs, const double *a, size_t lda,
                     const double *b, size_t ldb, double *x, size_t ldx);
int mat_gpu_posv_f32(void *stream, size_t n, size_t nrhs, const float *a, size_t lda,
                     const float *b, size_t ldb, float *x, size_t ldx);
int mat_gpu_posv_f64(void *stream, size_t n, size_t nrhs, const double *a, size_t lda,
                     const double *b, size_t ldb, double *x, size_t ldx);

/* Mixed-precision solvers with fp64 iterative refinement.
 *   passes: trailing-update multiply of the fp32 factorization (default
 *           6 for gesv, 8 for posv):
 *           0 = fp32 MFMA, 1/3/6 = BF16 MFMA with 1/3/6 split products,
 *           8 = FP8 e4m3 with per-128x128-tile scales (posv only; the
 *           trailing updates only, the panel solves use 3-pass BF16).
 *           Products below 2e9 multiply-adds run as fp32 MFMA.
 *   The residual b - A x (A in fp64) and x += d are the only fp64 steps.
 *   Stops at backward error max_c ||b_c - A x_c|| / (||A|| ||x_c|| + ||b_c||)
 *   <= tol (infinity norms).  posv with fp8: when the factor loses positive
 *   definiteness or refinement stalls, it refactors with 3-pass BF16 first.
 *   Falls back to the plain fp64 factorization when refinement stalls.
 *   Synchronous. */
typedef struct {
  int passes;   /* default 6 (gesv) / 8 (posv) */
  int max_iter; /* default 30 */
  double tol;   /* default 1e-14 */
  int leaf;     /* LU pivoting block width in [1, 64], default 32 */
  int pivoting; /* gesv only: 0 = partial pivoting, libmat's rule;
                   1 = recursive random butterfly transform + no pivoting
                   (faster factorization, but its growth can stall
                   refinement on ill-conditioned matrices); 2 = tournament
                   pivoting (CALU, default) for leaves of at most 32
                   columns and 65536 rows (partial pivoting elsewhere):
                   the workgroups meet once per level of a reduction tree
                   (three levels at 16384 rows) instead of once per
                   column */
} mat_gpu_ir_opts;
typedef struct {
  int iterations;        /* refinement steps taken */
  int fell_back;         /* 1: the plain fp64 factorization produced x */
  double backward_error; /* final normwise backward error */
  double factor_ms;      /* low-precision factorization wall time */
  double total_ms;       /* whole call wall time */
  int precision;         /* passes of the factorization that produced x
                            (posv: 8 = fp8, or 3 after the fp8 factor lost
                            definiteness or stalled; see fell_back) */
} mat_gpu_ir_status;
/* Fill opts with the defaults. */
void mat_gpu_ir_defaults(mat_gpu_ir_opts *opts, int spd);
int mat_gpu_gesv_mixed(void *stream, size_t n, size_t nrhs, const double *a, size_t lda,
                       const double *b, size_t ldb, double *x, size_t ldx,
                       const mat_gpu_ir_opts *opts, mat_gpu_ir_status *status);
int mat_gpu_posv_mixed(void *stream, size_t n, size_t nrhs, const double *a, size_t lda,
                       const double *b, size_t ldb, double *x, size_t ldx,
                       const mat_gpu_ir_opts *opts, mat_gpu_ir_status *status);

/* Batched small solves (n in [1, 64], one right-hand side each), fp32.
 * Matrix i: order n[i], kind[i] = 0 general (partial-pivot LU, mat_solve
 * semantics) or 1 SPD (Cholesky, mat_solve_spd semantics: info[i] = -1 and
 * x untouched when not positive definite); column-major at a + a_off[i];
 * right-hand side b + v_off[i], solution x + v_off[i].  n and kind are host
 * arrays; everything else is device memory.
 * A plan sorts the batch into size bins once and captures the launches (a
 * counter reset and one persistent kernel per size bin and kind) in a
 * hipGraph; mat_gpu_batch_run replays it (the data
 * behind the pointers may change between runs, the pointers and sizes may
 * not).  autotune != 0 times candidate plans on this batch and keeps the
 * fastest: the bins in use, the wavefronts per workgroup of every class,
 * the resident workgroups per CU and the class order. */
typedef struct mat_gpu_batch_plan mat_gpu_batch_plan;
int mat_gpu_batch_create_f32(void *stream, size_t count, const int32_t *n, const int8_t *kind,
                             const float *a, const int64_t *a_off, const float *b, float *x,
                             const int64_t *v_off, int32_t *info, int autotune,
                             mat_gpu_batch_plan **plan);
int mat_gpu_batch_run(void *stream, mat_gpu_batch_plan *plan);
int mat_gpu_batch_tuning(const mat_gpu_batch_plan *plan, int *wg_per_cu, int *order_mode,
                         double *tuned_ms);
/* The autotuner's shape choices: bin_mask bit b set = the bin of groups of
 * 8 << b lanes is in use (an order goes to the smallest bin in use that
 * holds it; bit 3 is always set), waves[8] = wavefronts per workgroup of
 * the class (bin b, kind k) at index 2b + k, and the time of the untuned
 * plan (0 for a plan built without autotuning). */
int mat_gpu_batch_tuning_shape(const mat_gpu_batch_plan *plan, int *bin_mask, int32_t *waves,
                               double *untuned_ms);
int mat_gpu_batch_destroy(mat_gpu_batch_plan *plan);
/* Plan + one ordinary launch, no graph; synchronous. */
int mat_gpu_batch_solve_f32(void *stream, size_t count, const int32_t *n, const int8_t *kind,
                            const float *a, const int64_t *a_off, const float *b, float *x,
                            const int64_t *v_off, int32_t *info);
/* mat_gpu_batch_solve_f32 that allocates nothing (no counterpart in libmat):
 * for callers that cannot allocate memory, such as serving loops that
 * forbid it.  Its scratch (the orders, the task queue and the per-class
 * counters) lives in one fixed array of about 6 MiB per GPU, held until the
 * process ends, so a call takes at most 262144 matrices (larger batches:
 * a negative code and a message).  Same kernels and results as
 * mat_gpu_batch_solve_f32.  Every call on one GPU shares that array: the
 * library serializes calls within a process (a lock held for the whole,
 * synchronous call); nothing stops two processes from sharing a GPU, each
 * with its own array.  Synchronous. */
int mat_gpu_batch_solve_noalloc_f32(void *stream, size_t count, const int32_t *n,
                                    const int8_t *kind, const float *a, const int64_t *a_off,
                                    const float *b, float *x, const int64_t *v_off,
                                    int32_t *info);

#ifdef __cplusplus
}
#endif
#endif /* MAT_GPU_H_ */
