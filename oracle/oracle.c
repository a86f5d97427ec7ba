/*
 * CPU oracle for the dense GEMM + factor-and-solve path of libmat.
 *
 * TEST INFRASTRUCTURE ONLY. This file is a clean-room restatement of the
 * x86 scalar code path of libmat (the path MAT_DISPATCH selects on x86 when
 * neither NEON nor AVX2 is enabled).  It exists so that the tests, the
 * smoke check and the CPU-baseline leg of the benchmark can check and time
 * the GPU library against the reference algorithm.  Nothing in the shipped
 * library links or calls it.
 *
 * Every routine repeats the reference's loop order and operation order so
 * that, built with -O2 -ffp-contract=off and no -march, it reproduces the
 * reference's outputs bit for bit.  That claim is pinned by
 * tests/test_oracle_vs_ref.py (against oracle/_ref, the reference compiled
 * from its own sources) and tests/test_oracle_golden.py (against committed
 * fixtures and the reference's own known-answer tests).
 *
 * Compiled twice: ORACLE_F64 undefined -> float flavour (mat_elem_t=float,
 * eps 1e-6f); ORACLE_F64 defined -> double flavour (eps 1e-9), mirroring
 * mat.h:165-181.
 *
 * Storage is column-major with leading dimension = rows (mat.h:387-393).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORACLE_F64
typedef double elem;
#define ORC_EPS 1e-9
#define ORC_ABS fabs
#define ORC_SQRT sqrt
#define ORC(name) orc_d_##name
#else
typedef float elem;
#define ORC_EPS 1e-6f
#define ORC_ABS fabsf
#define ORC_SQRT sqrtf
#define ORC(name) orc_s_##name
#endif

#define AT(p, ld, i, j) ((p)[(size_t)(j) * (ld) + (i)])

/* y *= alpha with the reference's special cases: alpha == 1 returns,
 * alpha == 0 zero-fills (mat.h:959-964, 993-1000). */
static void scal(elem *y, elem alpha, size_t n) {
  if (alpha == 1) return;
  if (alpha == 0) {
    memset(y, 0, n * sizeof(elem));
    return;
  }
  for (size_t i = 0; i < n; i++) y[i] *= alpha;
}

/* y += alpha * x (mat.h:908-914). */
static void axpy(elem *y, elem alpha, const elem *x, size_t n) {
  for (size_t i = 0; i < n; i++) y[i] += alpha * x[i];
}

/* Index of the first element whose magnitude equals max|x|
 * (mat.h:1195-1204 amax, 1245-1262 iamax). */
static size_t iamax(const elem *x, size_t n) {
  if (n == 0) return 0;
  elem m = ORC_ABS(x[0]);
  for (size_t i = 1; i < n; i++) {
    elem v = ORC_ABS(x[i]);
    if (v > m) m = v;
  }
  for (size_t i = 0; i < n; i++)
    if (ORC_ABS(x[i]) == m) return i;
  return 0;
}

/* C = alpha*A*B + beta*C, A MxK, B KxN (mat_gemm -> mat__gemm_scalar,
 * mat.h:2762-2796): beta-scale, then i-k-j loops with alpha folded into
 * A(i,k). */
void ORC(gemm)(size_t M, size_t K, size_t N, elem alpha, const elem *A,
               const elem *B, elem beta, elem *C) {
  scal(C, beta, M * N);
  for (size_t i = 0; i < M; i++) {
    for (size_t k = 0; k < K; k++) {
      elem aik = alpha * AT(A, M, i, k);
      for (size_t j = 0; j < N; j++)
        AT(C, M, i, j) = AT(C, M, i, j) + aik * AT(B, K, k, j);
    }
  }
}

/* Symmetric rank-k update on one triangle of the n x n matrix C:
 * C = alpha*A*A^T + beta*C with A n x k (mat_syrk -> mat__syrk_scalar,
 * mat.h:5641-5668, 5670-5677), or C = alpha*A^T*A + beta*C with A k x n
 * when trans (mat_syrk_t -> mat__syrk_t_scalar, mat.h:5698-5725,
 * 5727-5734).  Row i outer, then j over the triangle (lower: j <= i,
 * upper: j >= i); the k-sum runs in order, then C(i,j) = beta*C(i,j) +
 * alpha*sum.  The other triangle is not touched. */
void ORC(syrk)(size_t n, size_t k, elem alpha, const elem *A, elem beta,
               elem *C, int upper, int trans) {
  for (size_t i = 0; i < n; i++) {
    const size_t j0 = upper ? i : 0, j1 = upper ? n : i + 1;
    for (size_t j = j0; j < j1; j++) {
      elem sum = 0;
      for (size_t kk = 0; kk < k; kk++) {
        const elem a = trans ? AT(A, k, kk, i) : AT(A, n, i, kk);
        const elem b = trans ? AT(A, k, kk, j) : AT(A, n, j, kk);
        sum += a * b;
      }
      AT(C, n, i, j) = beta * AT(C, n, i, j) + alpha * sum;
    }
  }
}

/* Strided GEMM without transposes, scalar branch of mat__gemm_strided_
 * (mat.h:2719-2743): per-column beta-scale, then j-k-i with alpha folded
 * into B(k,j). */
static void gemm_strided_nn(elem *C, size_t ldc, elem alpha, const elem *A,
                            size_t lda, const elem *B, size_t ldb, size_t M,
                            size_t K, size_t N, elem beta) {
  for (size_t j = 0; j < N; j++) scal(&C[j * ldc], beta, M);
  for (size_t j = 0; j < N; j++) {
    for (size_t k = 0; k < K; k++) {
      elem ab = alpha * B[j * ldb + k];
      for (size_t i = 0; i < M; i++) C[j * ldc + i] += A[k * lda + i] * ab;
    }
  }
}

#define PLU_NB 24 /* MAT_PLU_BLOCK_SIZE, mat.h:4196 */

/* In-place blocked right-looking partial-pivot LU with delayed row swaps
 * (mat__plu_blocked_, mat.h:4201-4288).  p holds the row permutation
 * (p[i] = original row now at row i); returns the number of row swaps. */
static int plu_inplace(size_t n, elem *a, size_t *p) {
  int swaps = 0;
  size_t piv[PLU_NB];
  for (size_t kb = 0; kb < n; kb += PLU_NB) {
    size_t ke = kb + PLU_NB < n ? kb + PLU_NB : n;
    size_t bk = ke - kb;
    for (size_t k = kb; k < ke && k < n - 1; k++) {
      elem *ck = &a[k * n];
      size_t pr = k + iamax(&ck[k], n - k);
      piv[k - kb] = pr;
      if (pr != k) {
        for (size_t j = 0; j < ke; j++) {
          elem t = a[j * n + k];
          a[j * n + k] = a[j * n + pr];
          a[j * n + pr] = t;
        }
        size_t t = p[k];
        p[k] = p[pr];
        p[pr] = t;
        swaps++;
      }
      if (ORC_ABS(ck[k]) < ORC_EPS) continue; /* near-zero pivot: skip */
      elem inv = 1.0f / ck[k];
      scal(&ck[k + 1], inv, n - (k + 1));
      for (size_t j = k + 1; j < ke; j++)
        axpy(&a[j * n + k + 1], -a[j * n + k], &ck[k + 1], n - (k + 1));
    }
    if (ke < n) {
      for (size_t k = kb; k < ke && k < n - 1; k++) {
        size_t pr = piv[k - kb];
        if (pr != k) {
          for (size_t j = ke; j < n; j++) {
            elem t = a[j * n + k];
            a[j * n + k] = a[j * n + pr];
            a[j * n + pr] = t;
          }
        }
      }
      for (size_t k = kb; k < ke; k++)
        for (size_t j = ke; j < n; j++)
          axpy(&a[j * n + k + 1], -a[j * n + k], &a[k * n + k + 1],
               ke - (k + 1));
      gemm_strided_nn(&a[ke * n + ke], n, -1.0f, &a[kb * n + ke], n,
                      &a[ke * n + kb], n, n - ke, bk, n - ke, 1.0f);
    }
  }
  return swaps;
}

/* Split packed LU into unit-lower L and upper U (mat.h:4456-4470). */
static void split_lu(size_t n, const elem *m, elem *L, elem *U) {
  memset(L, 0, n * n * sizeof(elem));
  memset(U, 0, n * n * sizeof(elem));
  for (size_t i = 0; i < n; i++) L[i * n + i] = 1;
  for (size_t j = 0; j < n; j++) {
    for (size_t i = j + 1; i < n; i++) L[j * n + i] = m[j * n + i];
    for (size_t i = 0; i <= j; i++) U[j * n + i] = m[j * n + i];
  }
}

/* mat_plu (mat.h:4436-4474): PA = LU, returns row-swap count. */
int ORC(plu)(size_t n, const elem *A, elem *L, elem *U, size_t *p) {
  elem *m = (elem *)malloc(n * n * sizeof(elem));
  memcpy(m, A, n * n * sizeof(elem));
  for (size_t i = 0; i < n; i++) p[i] = i;
  int s = plu_inplace(n, m, p);
  split_lu(n, m, L, U);
  free(m);
  return s;
}

/* Packed form of mat_plu for tests that want the in-place factor. */
int ORC(plu_packed)(size_t n, elem *a, size_t *p) {
  for (size_t i = 0; i < n; i++) p[i] = i;
  return plu_inplace(n, a, p);
}

/* mat_lu, x86 scalar branch (mat.h:4476-4578): full pivoting PAQ = LU.
 * Pivot scan is row-major (i outer, j inner), strict '>' starting from
 * |A(k,k)|; a step is skipped when the max is below eps; rows are swapped
 * before columns.  Returns row swaps + column swaps. */
int ORC(lu)(size_t n, const elem *A, elem *L, elem *U, size_t *p,
            size_t *q) {
  elem *d = (elem *)malloc(n * n * sizeof(elem));
  memcpy(d, A, n * n * sizeof(elem));
  for (size_t i = 0; i < n; i++) p[i] = q[i] = i;
  int swaps = 0;
  for (size_t k = 0; k + 1 < n; k++) {
    size_t pr = k, pc = k;
    elem mx = ORC_ABS(d[k * n + k]);
    for (size_t i = k; i < n; i++)
      for (size_t j = k; j < n; j++) {
        elem v = ORC_ABS(d[j * n + i]);
        if (v > mx) {
          mx = v;
          pr = i;
          pc = j;
        }
      }
    if (mx < ORC_EPS) continue;
    if (pr != k) {
      for (size_t j = 0; j < n; j++) {
        elem t = d[j * n + k];
        d[j * n + k] = d[j * n + pr];
        d[j * n + pr] = t;
      }
      size_t t = p[k];
      p[k] = p[pr];
      p[pr] = t;
      swaps++;
    }
    if (pc != k) {
      for (size_t i = 0; i < n; i++) {
        elem t = d[k * n + i];
        d[k * n + i] = d[pc * n + i];
        d[pc * n + i] = t;
      }
      size_t t = q[k];
      q[k] = q[pc];
      q[pc] = t;
      swaps++;
    }
    elem inv = 1.0f / d[k * n + k];
    for (size_t i = k + 1; i < n; i++) {
      elem l = d[k * n + i] * inv;
      d[k * n + i] = l;
      for (size_t j = k + 1; j < n; j++) d[j * n + i] -= l * d[j * n + k];
    }
  }
  split_lu(n, d, L, U);
  free(d);
  return swaps;
}

/* mat_chol (mat.h:6062-6095) with the x86 kernel mat__chol_scalar_generic_
 * (mat.h:6031-6060): copy the lower triangle, zero the upper, then
 * left-looking column Cholesky; -1 at the first non-positive pivot. */
int ORC(chol)(size_t n, const elem *A, elem *L) {
  for (size_t j = 0; j < n; j++) {
    for (size_t i = 0; i < j; i++) L[j * n + i] = 0;
    memcpy(&L[j * n + j], &A[j * n + j], (n - j) * sizeof(elem));
  }
  for (size_t j = 0; j < n; j++) {
    elem s = 0;
    for (size_t k = 0; k < j; k++) {
      elem v = AT(L, n, j, k);
      s += v * v;
    }
    elem dg = AT(L, n, j, j) - s;
    if (dg <= 0) return -1;
    elem ljj = ORC_SQRT(dg);
    AT(L, n, j, j) = ljj;
    elem inv = 1.0f / ljj;
    for (size_t i = j + 1; i < n; i++) {
      elem dot = 0;
      for (size_t k = 0; k < j; k++) dot += AT(L, n, i, k) * AT(L, n, j, k);
      AT(L, n, i, j) = (AT(L, n, i, j) - dot) * inv;
    }
  }
  return 0;
}

/* Lx = b, non-unit lower (mat__solve_tril_scalar, mat.h:4584-4598). */
void ORC(solve_tril)(size_t n, const elem *L, const elem *b, elem *x) {
  for (size_t i = 0; i < n; i++) {
    elem dot = 0;
    for (size_t j = 0; j < i; j++) dot += AT(L, n, i, j) * x[j];
    x[i] = (b[i] - dot) / AT(L, n, i, i);
  }
}

/* Lx = b, unit lower (mat__solve_tril_unit_scalar, mat.h:4638-4652). */
void ORC(solve_tril_unit)(size_t n, const elem *L, const elem *b, elem *x) {
  for (size_t i = 0; i < n; i++) {
    elem dot = 0;
    for (size_t j = 0; j < i; j++) dot += AT(L, n, i, j) * x[j];
    x[i] = b[i] - dot;
  }
}

/* Ux = b, upper (mat__solve_triu_scalar, mat.h:4722-4736). */
void ORC(solve_triu)(size_t n, const elem *U, const elem *b, elem *x) {
  for (size_t i = n; i-- > 0;) {
    elem s = b[i];
    for (size_t j = i + 1; j < n; j++) s -= AT(U, n, i, j) * x[j];
    x[i] = s / AT(U, n, i, i);
  }
}

/* L^T x = b via column dots (mat__solve_trilt_scalar, mat.h:4773-4790). */
void ORC(solve_trilt)(size_t n, const elem *L, const elem *b, elem *x) {
  for (size_t j = n; j-- > 0;) {
    const elem *c = &L[j * n];
    elem dot = 0;
    for (size_t k = j + 1; k < n; k++) dot += c[k] * x[k];
    x[j] = (b[j] - dot) / c[j];
  }
}

/* mat_solve (mat.h:4807-4837): PLU, permute b, unit-lower then upper
 * substitution.  One right-hand side. */
void ORC(solve)(size_t n, const elem *A, const elem *b, elem *x) {
  elem *L = (elem *)malloc(n * n * sizeof(elem));
  elem *U = (elem *)malloc(n * n * sizeof(elem));
  size_t *p = (size_t *)malloc(n * sizeof(size_t));
  elem *pb = (elem *)malloc(n * sizeof(elem));
  elem *z = (elem *)malloc(n * sizeof(elem));
  ORC(plu)(n, A, L, U, p);
  for (size_t i = 0; i < n; i++) pb[i] = b[p[i]];
  ORC(solve_tril_unit)(n, L, pb, z);
  ORC(solve_triu)(n, U, z, x);
  free(L);
  free(U);
  free(p);
  free(pb);
  free(z);
}

/* mat_solve_spd (mat.h:4839-4871): Cholesky, then L y = b, L^T x = y.
 * Returns -1 (x untouched) when A is not positive definite. */
int ORC(solve_spd)(size_t n, const elem *A, const elem *b, elem *x) {
  elem *L = (elem *)malloc(n * n * sizeof(elem));
  elem *y = (elem *)calloc(n, sizeof(elem));
  int r = ORC(chol)(n, A, L);
  if (r != 0) {
    free(L);
    free(y);
    return -1;
  }
  ORC(solve_tril)(n, L, b, y);
  ORC(solve_trilt)(n, L, y, x);
  free(L);
  free(y);
  return 0;
}

/* Factor once, then k right-hand sides through the two substitutions,
 * which is what mat_solve does internally per call (mat.h:4823-4830).
 * B and X are n x nrhs, column-major. Used as the multi-RHS CPU baseline. */
void ORC(solve_multi)(size_t n, size_t nrhs, const elem *A, const elem *B,
                      elem *X) {
  elem *L = (elem *)malloc(n * n * sizeof(elem));
  elem *U = (elem *)malloc(n * n * sizeof(elem));
  size_t *p = (size_t *)malloc(n * sizeof(size_t));
  elem *pb = (elem *)malloc(n * sizeof(elem));
  elem *z = (elem *)malloc(n * sizeof(elem));
  ORC(plu)(n, A, L, U, p);
  for (size_t c = 0; c < nrhs; c++) {
    for (size_t i = 0; i < n; i++) pb[i] = B[c * n + p[i]];
    ORC(solve_tril_unit)(n, L, pb, z);
    ORC(solve_triu)(n, U, z, &X[c * n]);
  }
  free(L);
  free(U);
  free(p);
  free(pb);
  free(z);
}

/* Cholesky once, then per right-hand side L y = b and L^T x = y: what
 * mat_solve_spd does internally per call (mat.h:4854-4865).  B and X are
 * n x nrhs, column-major.  Returns mat_chol's result; X is untouched when
 * it is -1.  Used as the multi-RHS Cholesky CPU baseline. */
int ORC(chol_solve_multi)(size_t n, size_t nrhs, const elem *A, const elem *B,
                          elem *X) {
  elem *L = (elem *)malloc(n * n * sizeof(elem));
  elem *y = (elem *)calloc(n, sizeof(elem));
  int r = ORC(chol)(n, A, L);
  if (r == 0) {
    for (size_t c = 0; c < nrhs; c++) {
      ORC(solve_tril)(n, L, &B[c * n], y);
      ORC(solve_trilt)(n, L, y, &X[c * n]);
    }
  }
  free(L);
  free(y);
  return r;
}

/* A batch of independent small systems, one mat_solve (kind 0) or
 * mat_solve_spd (kind 1) call per matrix, in order: the per-matrix CPU
 * loop the batched GPU solver replaces.  Matrix i is n[i] x n[i] at
 * A + a_off[i]; its right-hand side and solution at b / x + v_off[i];
 * info[i] gets mat_solve_spd's result (0 for kind 0). */
void ORC(batch_solve)(size_t count, const int *n, const signed char *kind,
                      const elem *A, const long long *a_off, const elem *b,
                      elem *x, const long long *v_off, int *info) {
  for (size_t i = 0; i < count; i++) {
    const size_t ni = (size_t)n[i];
    if (kind[i] == 0) {
      ORC(solve)(ni, A + a_off[i], b + v_off[i], x + v_off[i]);
      info[i] = 0;
    } else {
      info[i] = ORC(solve_spd)(ni, A + a_off[i], b + v_off[i], x + v_off[i]);
    }
  }
}

/* mat_inv (mat.h:6166-6225): blocked PLU of a copy, out = the permuted
 * identity (column i holds a one at row p[i]), then in place the
 * unit-lower forward elimination column j by column j over every column
 * of out (zero entries skipped), then the upper back substitution from
 * the last column, each row scaled by 1/U(j,j) before it updates the rows
 * above. */
void ORC(inv)(size_t n, const elem *A, elem *out) {
  elem *m = (elem *)malloc(n * n * sizeof(elem));
  size_t *p = (size_t *)malloc(n * sizeof(size_t));
  memcpy(m, A, n * n * sizeof(elem));
  for (size_t i = 0; i < n; i++) p[i] = i;
  plu_inplace(n, m, p);
  memset(out, 0, n * n * sizeof(elem));
  for (size_t i = 0; i < n; i++) out[i * n + p[i]] = 1.0f;
  for (size_t j = 0; j < n; j++) {
    const elem *Lj = &m[j * n];
    for (size_t col = 0; col < n; col++) {
      elem *oc = &out[col * n];
      const elem ojc = oc[j];
      if (ojc == 0) continue;
      for (size_t i = j + 1; i < n; i++) oc[i] -= ojc * Lj[i];
    }
  }
  for (size_t jj = n; jj > 0; jj--) {
    const size_t j = jj - 1;
    const elem *Uj = &m[j * n];
    const elem dinv = 1.0f / Uj[j];
    for (size_t col = 0; col < n; col++) {
      elem *oc = &out[col * n];
      oc[j] *= dinv;
      const elem ojc = oc[j];
      for (size_t i = 0; i < j; i++) oc[i] -= ojc * Uj[i];
    }
  }
  free(m);
  free(p);
}

/* mat_det (mat.h:6227-6249): (-1)^swaps * prod diag(U). */
elem ORC(det)(size_t n, const elem *A) {
  elem *L = (elem *)malloc(n * n * sizeof(elem));
  elem *U = (elem *)malloc(n * n * sizeof(elem));
  size_t *p = (size_t *)malloc(n * sizeof(size_t));
  int s = ORC(plu)(n, A, L, U, p);
  elem d = (s % 2 == 0) ? 1 : -1;
  for (size_t i = 0; i < n; i++) d *= U[i * n + i];
  free(L);
  free(U);
  free(p);
  return d;
}

/* ---------------------------------------------------------------- QR --- */

/* Strided GEMM with transpose flags, scalar branch of mat__gemm_strided_
 * (mat.h:2719-2743): per-column beta-scale (mat__scal_kernel's special
 * cases), then j-k-i with alpha folded into op(B)(k,j). */
static void gemm_strided(elem *C, size_t ldc, elem alpha, const elem *A,
                         size_t lda, int ta, const elem *B, size_t ldb, int tb,
                         size_t M, size_t K, size_t N, elem beta) {
  for (size_t j = 0; j < N; j++) scal(&C[j * ldc], beta, M);
  for (size_t j = 0; j < N; j++) {
    for (size_t k = 0; k < K; k++) {
      elem ab = alpha * (tb ? B[k * ldb + j] : B[j * ldb + k]);
      for (size_t i = 0; i < M; i++)
        C[j * ldc + i] += (ta ? A[i * lda + k] : A[k * lda + i]) * ab;
    }
  }
}

static elem dot(const elem *a, const elem *b, size_t n) {
  elem r = 0;
  for (size_t i = 0; i < n; i++) r += a[i] * b[i];
  return r;
}

/* Householder vector of x (mat_householder, mat.h:3634-3716): v[0] = 1,
 * v[i] = x[i] / (x0 - beta), beta = -sign(x0) ||x||, tau = -v0 / beta;
 * ||x|| < eps gives v = 0, tau = 0.  v may alias x. */
static void householder(elem *v, elem *tau, const elem *x, size_t n) {
  if (n == 0) {
    *tau = 0;
    return;
  }
  elem norm_sq = 0;
  for (size_t i = 0; i < n; i++) norm_sq += x[i] * x[i];
  elem norm_x = ORC_SQRT(norm_sq);
  if (norm_x < ORC_EPS) {
    *tau = 0;
    for (size_t i = 0; i < n; i++) v[i] = 0;
    return;
  }
  elem x0 = x[0];
  elem beta = (x0 >= 0) ? -norm_x : norm_x;
  elem v0 = x0 - beta;
  elem v0_inv = 1 / v0;
  v[0] = 1;
  for (size_t i = 1; i < n; i++) v[i] = x[i] * v0_inv;
  *tau = -v0 / beta;
}

#define QR_NB 16        /* MAT_QR_BLOCK_SIZE, mat.h:3762 */
#define QR_THRESHOLD 64 /* MAT_QR_BLOCK_THRESHOLD, mat.h:3766 */

/* Compact-WY factor T of k reflectors (mat__qr_build_T_, mat.h:3772-3827):
 * YtY(p,j) = Y(:,p).Y(:,j) for p < j, T(j,j) = tau_j,
 * T(i,j) = -tau_j * sum_{p=i}^{j-1} T(i,p) YtY(p,j). */
static void build_T(elem *T, size_t ldt, const elem *Y, size_t ldy,
                    const elem *tau, size_t m, size_t k) {
  for (size_t i = 0; i < k * k; i++) T[i] = 0;
  elem *YtY = (elem *)malloc(k * k * sizeof(elem));
  for (size_t j = 1; j < k; j++)
    for (size_t p = 0; p < j; p++) YtY[j * k + p] = dot(&Y[p * ldy], &Y[j * ldy], m);
  for (size_t j = 0; j < k; j++) {
    T[j * ldt + j] = tau[j];
    for (size_t i = 0; i < j; i++) {
      elem sum = 0;
      for (size_t p = i; p < j; p++) sum += T[p * ldt + i] * YtY[j * k + p];
      T[j * ldt + i] = -tau[j] * sum;
    }
  }
  free(YtY);
}

/* Blocked and unblocked Householder QR (mat_qr, mat.h:3907-4063, with the
 * block applies of mat.h:3846-3903; mat_qr_r, mat.h:4066-4192, when Q is
 * NULL).  A is m x n (m >= n), Q m x m, R m x n, all column-major.  R
 * keeps what the reflectors leave below its diagonal (rounding residue),
 * as the reference does. */
static void qr_impl(size_t m, size_t n, const elem *A, elem *Q, elem *R) {
  memcpy(R, A, m * n * sizeof(elem));
  if (Q) {
    memset(Q, 0, m * m * sizeof(elem));
    for (size_t i = 0; i < m; i++) Q[i * m + i] = 1;
  }
  if (n >= QR_THRESHOLD) {
    const size_t nb = QR_NB;
    elem *tau = (elem *)malloc(nb * sizeof(elem));
    elem *Y = (elem *)malloc(m * nb * sizeof(elem));
    elem *T = (elem *)malloc(nb * nb * sizeof(elem));
    elem *x = (elem *)malloc(m * sizeof(elem));
    elem *Wmk = (elem *)malloc(m * nb * sizeof(elem));
    elem *Wmk2 = (elem *)malloc(m * nb * sizeof(elem));
    elem *Wkn = (elem *)malloc(nb * n * sizeof(elem));
    elem *Wkn2 = (elem *)malloc(nb * n * sizeof(elem));
    for (size_t jb = 0; jb < n; jb += nb) {
      const size_t kb = (jb + nb <= n) ? nb : n - jb;
      const size_t len = m - jb;
      for (size_t j = 0; j < kb; j++) {
        const size_t col = jb + j, vlen = m - col;
        memcpy(x, &R[col * m + col], vlen * sizeof(elem));
        householder(x, &tau[j], x, vlen);
        for (size_t i = 0; i < j; i++) Y[j * len + i] = 0;
        for (size_t i = 0; i < vlen; i++) Y[j * len + j + i] = x[i];
        if (tau[j] != 0) {
          for (size_t c = col; c < jb + kb; c++) {
            elem *rc = &R[c * m + col];
            elem w = dot(x, rc, vlen) * tau[j];
            axpy(rc, -w, x, vlen);
          }
        }
      }
      build_T(T, kb, Y, len, tau, len, kb);
      if (jb + kb < n) {
        /* R(jb:, jb+kb:) -= Y T^T (Y^T R(jb:, jb+kb:)), mat.h:3846-3869 */
        const size_t tc = n - (jb + kb);
        elem *Rs = &R[(jb + kb) * m + jb];
        gemm_strided(Wkn, kb, 1, Y, len, 1, Rs, m, 0, kb, len, tc, 0);
        gemm_strided(Wkn2, kb, 1, T, kb, 1, Wkn, kb, 0, kb, kb, tc, 0);
        gemm_strided(Rs, m, -1, Y, len, 0, Wkn2, kb, 0, len, kb, tc, 1);
      }
      if (Q) {
        /* Q(:, jb:) -= (Q(:, jb:) Y) T Y^T, mat.h:3873-3903 */
        elem *Qt = &Q[jb * m];
        gemm_strided(Wmk, m, 1, Qt, m, 0, Y, len, 0, m, len, kb, 0);
        gemm_strided(Wmk2, m, 1, Wmk, m, 0, T, kb, 0, m, kb, kb, 0);
        gemm_strided(Qt, m, -1, Wmk2, m, 0, Y, len, 1, m, kb, len, 1);
      }
    }
    free(tau);
    free(Y);
    free(T);
    free(x);
    free(Wmk);
    free(Wmk2);
    free(Wkn);
    free(Wkn2);
    return;
  }
  elem *x = (elem *)malloc(m * sizeof(elem));
  elem *v = (elem *)malloc(m * sizeof(elem));
  elem *u = (elem *)malloc(m * sizeof(elem));
  for (size_t j = 0; j < n; j++) {
    const size_t len = m - j;
    elem tau;
    memcpy(x, &R[j * m + j], len * sizeof(elem));
    householder(v, &tau, x, len);
    if (tau == 0) continue;
    for (size_t k = j; k < n; k++) {
      elem *rk = &R[k * m + j];
      elem w = tau * dot(v, rk, len);
      axpy(rk, -w, v, len);
    }
    if (Q) {
      /* Q(:, j:) = Q(:, j:) - tau (Q(:, j:) v) v^T, column by column */
      memset(u, 0, m * sizeof(elem));
      for (size_t jj = 0; jj < len; jj++) axpy(u, v[jj], &Q[(j + jj) * m], m);
      for (size_t jj = 0; jj < len; jj++) axpy(&Q[(j + jj) * m], -tau * v[jj], u, m);
    }
  }
  free(x);
  free(v);
  free(u);
}

void ORC(qr)(size_t m, size_t n, const elem *A, elem *Q, elem *R) { qr_impl(m, n, A, Q, R); }

void ORC(qr_r)(size_t m, size_t n, const elem *A, elem *R) { qr_impl(m, n, A, NULL, R); }
