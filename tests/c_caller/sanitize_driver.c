/* Runs the CPU oracle (oracle/oracle.c) and the library's host helpers
 * (libmat_amd/csrc/mat_host.cpp) over many shapes, for a build with
 * AddressSanitizer and UBSan (tests/test_sanitizers.py): any out-of-bounds
 * access, leak or undefined behaviour aborts the run.  Build with the same
 * precision flags as the two sources (ORACLE_F64 and MAT_DOUBLE_PRECISION
 * together for doubles). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mat_gpu.h"

typedef mat_elem_t elem;
#ifdef ORACLE_F64
#define ORC(name) orc_d_##name
#else
#define ORC(name) orc_s_##name
#endif

void ORC(gemm)(size_t, size_t, size_t, elem, const elem *, const elem *, elem, elem *);
void ORC(syrk)(size_t, size_t, elem, const elem *, elem, elem *, int, int);
int ORC(plu)(size_t, const elem *, elem *, elem *, size_t *);
int ORC(lu)(size_t, const elem *, elem *, elem *, size_t *, size_t *);
int ORC(chol)(size_t, const elem *, elem *);
void ORC(solve_tril)(size_t, const elem *, const elem *, elem *);
void ORC(solve_tril_unit)(size_t, const elem *, const elem *, elem *);
void ORC(solve_triu)(size_t, const elem *, const elem *, elem *);
void ORC(solve_trilt)(size_t, const elem *, const elem *, elem *);
void ORC(solve)(size_t, const elem *, const elem *, elem *);
int ORC(solve_spd)(size_t, const elem *, const elem *, elem *);
void ORC(solve_multi)(size_t, size_t, const elem *, const elem *, elem *);
elem ORC(det)(size_t, const elem *);
void ORC(qr)(size_t, size_t, const elem *, elem *, elem *);
void ORC(qr_r)(size_t, size_t, const elem *, elem *);
void ORC(inv)(size_t, const elem *, elem *);
int ORC(chol_solve_multi)(size_t, size_t, const elem *, const elem *, elem *);
void ORC(batch_solve)(size_t, const int *, const signed char *, const elem *, const long long *,
                      const elem *, elem *, const long long *, int *);

static unsigned long long rng = 88172645463325252ull;
static elem uni(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (elem)((double)(rng >> 11) / 9007199254740992.0 * 2.0 - 1.0);
}
static elem *fill(size_t count) {
  elem *p = (elem *)malloc((count ? count : 1) * sizeof(elem));
  for (size_t i = 0; i < count; ++i) p[i] = uni();
  return p;
}

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, #c);     \
      ++failures;                                                 \
    }                                                             \
  } while (0)

static void oracle_shapes(size_t n) {
  const size_t k = n + 3;
  elem *A = fill(n * n), *B = fill(n * k), *C = fill(n * k), *b = fill(n), *x = fill(n);
  ORC(gemm)(n, n, k, (elem)0.5, A, B, (elem)2, C);
  elem *Ck = fill(n * n), *Ak = fill(n * k);
  ORC(syrk)(n, k, (elem)1, Ak, (elem)0.5, Ck, 0, 0);
  ORC(syrk)(n, k, (elem)1, Ak, (elem)0.5, Ck, 1, 1);
  elem *L = fill(n * n), *U = fill(n * n), *Q = fill(n * n), *R = fill(n * n);
  size_t *p = (size_t *)malloc(n * sizeof(size_t)), *q = (size_t *)malloc(n * sizeof(size_t));
  const int swaps = ORC(plu)(n, A, L, U, p);
  CHECK(swaps >= 0 && (size_t)swaps < n + 1);
  /* PA = LU */
  double err = 0, nrm = 0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      double s = 0;
      for (size_t t = 0; t < n; ++t) s += (double)L[t * n + i] * U[j * n + t];
      const double a = A[j * n + p[i]];
      err += (s - a) * (s - a);
      nrm += a * a;
    }
  CHECK(sqrt(err / nrm) < 1e-3);
  ORC(lu)(n, A, L, U, p, q);
  /* an SPD matrix: A A^T + n I */
  elem *S = (elem *)malloc(n * n * sizeof(elem));
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      double s = i == j ? (double)n : 0.0;
      for (size_t t = 0; t < n; ++t) s += (double)A[t * n + i] * A[t * n + j];
      S[j * n + i] = (elem)s;
    }
  CHECK(ORC(chol)(n, S, L) == 0);
  ORC(solve_tril)(n, L, b, x);
  ORC(solve_tril_unit)(n, L, b, x);
  ORC(solve_triu)(n, S, b, x);
  ORC(solve_trilt)(n, L, b, x);
  ORC(solve)(n, A, b, x);
  CHECK(ORC(solve_spd)(n, S, b, x) == 0);
  ORC(solve_multi)(n, 3, A, C, B);
  CHECK(ORC(chol_solve_multi)(n, 3, S, C, B) == 0);
  (void)ORC(det)(n, A);
  ORC(qr)(n, n, A, Q, R);
  ORC(qr_r)(n, n, A, R);
  ORC(inv)(n, A, Q);
  free(A), free(B), free(C), free(b), free(x), free(Ck), free(Ak), free(L), free(U), free(Q);
  free(R), free(p), free(q), free(S);
}

static void tall_qr(size_t m, size_t n) {
  elem *A = fill(m * n), *Q = fill(m * m), *R = fill(m * n);
  ORC(qr)(m, n, A, Q, R);
  ORC(qr_r)(m, n, A, R);
  free(A), free(Q), free(R);
}

static void batch(void) {
  enum { COUNT = 40 };
  int n[COUNT];
  signed char kind[COUNT];
  long long aoff[COUNT], voff[COUNT], na = 0, nv = 0;
  for (int i = 0; i < COUNT; ++i) {
    n[i] = 1 + (i * 7) % 64;
    kind[i] = (signed char)(i % 2);
    aoff[i] = na;
    voff[i] = nv;
    na += (long long)n[i] * n[i];
    nv += n[i];
  }
  elem *A = fill((size_t)na), *b = fill((size_t)nv), *x = fill((size_t)nv);
  for (int i = 0; i < COUNT; ++i) /* SPD members: strongly diagonal */
    for (int d = 0; d < n[i]; ++d) A[aoff[i] + (long long)d * n[i] + d] += (elem)(2 * n[i]);
  for (int i = 0; i < COUNT; ++i)
    if (kind[i])
      for (int r = 0; r < n[i]; ++r)
        for (int c = 0; c < r; ++c)
          A[aoff[i] + (long long)c * n[i] + r] = A[aoff[i] + (long long)r * n[i] + c];
  int info[COUNT];
  ORC(batch_solve)(COUNT, n, kind, A, aoff, b, x, voff, info);
  for (int i = 0; i < COUNT; ++i) CHECK(info[i] == 0);
  free(A), free(b), free(x);
}

static void host_helpers(void) {
  const elem v[6] = {1, 2, 3, 4, 5, 6};
  Mat *m = mat_from(2, 3, v);
  CHECK(mat_at(m, 1, 2) == 6 && m->data[1] == 4);
  mat_set_at(m, 0, 0, 9);
  Mat *c = mat_rdeep_copy(m);
  CHECK(mat_equals(m, c));
  Mat *d = mat_mat(2, 3);
  mat_deep_copy(d, m);
  CHECK(mat_equals_tol(d, c, (elem)0));
  mat_init(d, v);
  CHECK(!mat_equals(d, c));
  Mat *e = mat_reye(4), *z = mat_zeros(4, 4), *em = mat_empty(3, 3);
  mat_eye(z);
  CHECK(mat_equals(e, z));
  Vec *w = mat_vec_from(3, v), *w0 = mat_vec(3);
  CHECK(w->rows == 3 && w->cols == 1 && w0->data[2] == 0);
  Perm *p = mat_perm(4);
  mat_perm_identity(p);
  p->data[0] = 2, p->data[2] = 0;
  Mat *pm = mat_perm_mat(p);
  CHECK(mat_at(pm, 0, 2) == 1 && mat_at(pm, 2, 0) == 1);
  free(em);
  mat_free_mat(m), mat_free_mat(c), mat_free_mat(d), mat_free_mat(e), mat_free_mat(z);
  mat_free_mat(w), mat_free_mat(w0), mat_free_mat(pm);
  mat_free_perm(p);
}

int main(void) {
  const size_t sizes[] = {1, 2, 3, 5, 23, 24, 25, 33, 48, 64, 65, 97};
  for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; ++i) oracle_shapes(sizes[i]);
  tall_qr(10, 7);
  tall_qr(100, 70);
  batch();
  host_helpers();
  printf("sanitize_driver: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
