/* A plain C program written against libmat's API (mat.h), linked to the
 * GPU library instead of libmat's implementation: it is what a libmat user
 * compiles when switching to libmatgpu_f32.so / libmatgpu_f64.so.
 *
 * Build against the reference header (no MAT_IMPLEMENTATION, so mat.h only
 * declares the functions, mat.h:45-47, 359-361):
 *     cc -DMAT_HEADER='"mat.h"' -I<libmat> libmat_caller.c -l:libmatgpu_f32.so
 * or against this project's copy of the declarations:
 *     cc -DMAT_HEADER='"mat_gpu.h"' -Iinclude libmat_caller.c -l:libmatgpu_f32.so
 * Add -DMAT_DOUBLE_PRECISION and link libmatgpu_f64.so for doubles.
 *
 * Run: ./libmat_caller [threads]  -- known answers of libmat's own tests
 * (tests/test_mat_mul.c, test_mat_solve.c, test_mat_chol.c,
 * test_mat_solve_spd.c, test_mat_plu.c, test_mat_det.c), then `threads` host
 * threads (default 2) that each factor and solve their own random systems at
 * the same time and check the residuals.  Exit status 0 when every check
 * passes. */
#include MAT_HEADER

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

static int failures = 0;

#define CHECK(cond)                                                    \
  do {                                                                 \
    if (!(cond)) {                                                     \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, \
              #cond);                                                  \
      ++failures;                                                      \
    }                                                                  \
  } while (0)

static int near(double a, double b, double tol) { return fabs(a - b) <= tol; }

/* Every libmat symbol the library provides, referenced so that the link
 * fails if one is missing or misnamed. */
typedef void (*fn_t)(void);
static const fn_t linked[] = {
    (fn_t)mat_mul, (fn_t)mat_rmul, (fn_t)mat_gemm, (fn_t)mat_lu, (fn_t)mat_plu, (fn_t)mat_chol,
    (fn_t)mat_solve, (fn_t)mat_solve_spd, (fn_t)mat_solve_tril, (fn_t)mat_solve_tril_unit,
    (fn_t)mat_solve_triu, (fn_t)mat_solve_trilt, (fn_t)mat_det, (fn_t)mat_inv, (fn_t)mat_syrk,
    (fn_t)mat_syrk_t, (fn_t)mat_qr, (fn_t)mat_qr_r, (fn_t)mat_empty, (fn_t)mat_mat,
    (fn_t)mat_from, (fn_t)mat_init, (fn_t)mat_free_mat, (fn_t)mat_zeros, (fn_t)mat_eye,
    (fn_t)mat_reye, (fn_t)mat_vec, (fn_t)mat_vec_from, (fn_t)mat_perm, (fn_t)mat_free_perm,
    (fn_t)mat_perm_identity, (fn_t)mat_deep_copy, (fn_t)mat_rdeep_copy, (fn_t)mat_at,
    (fn_t)mat_set_at, (fn_t)mat_equals, (fn_t)mat_equals_tol};

static void triangular_and_full_pivot(void) {
  Mat *L = mat_from(2, 2, (mat_elem_t[]){2, 0, 1, 2});
  Mat *U = mat_from(2, 2, (mat_elem_t[]){2, 1, 0, 2});
  Mat *L1 = mat_from(2, 2, (mat_elem_t[]){1, 0, 3, 1});
  Vec *x = mat_vec(2);
  Vec *b = mat_vec_from(2, (mat_elem_t[]){2, 5});
  mat_solve_tril(x, L, b); /* 2 x0 = 2, x0 + 2 x1 = 5 */
  CHECK(near(x->data[0], 1, 1e-6) && near(x->data[1], 2, 1e-6));
  Vec *b4 = mat_vec_from(2, (mat_elem_t[]){4, 4});
  mat_solve_triu(x, U, b4); /* 2 x0 + x1 = 4, 2 x1 = 4 */
  CHECK(near(x->data[0], 1, 1e-6) && near(x->data[1], 2, 1e-6));
  mat_solve_trilt(x, L, b4); /* L^T = U */
  CHECK(near(x->data[0], 1, 1e-6) && near(x->data[1], 2, 1e-6));
  Vec *b1 = mat_vec_from(2, (mat_elem_t[]){1, 5});
  mat_solve_tril_unit(x, L1, b1); /* x0 = 1, 3 x0 + x1 = 5 */
  CHECK(near(x->data[0], 1, 1e-6) && near(x->data[1], 2, 1e-6));
  /* full pivoting on [1 2; 3 4]: the largest entry 4 is pivoted first */
  Mat *A = mat_from(2, 2, (mat_elem_t[]){1, 2, 3, 4});
  Mat *Lf = mat_mat(2, 2), *Uf = mat_mat(2, 2);
  Perm *p = mat_perm(2), *q = mat_perm(2);
  const int swaps = mat_lu(A, Lf, Uf, p, q);
  CHECK(p->data[0] == 1 && q->data[0] == 1 && swaps == 2);
  CHECK(near(mat_at(Uf, 0, 0), 4, 1e-6));
  for (size_t i = 0; i < 2; ++i)
    for (size_t j = 0; j < 2; ++j) {
      double lu = 0;
      for (size_t k = 0; k < 2; ++k) lu += (double)mat_at(Lf, i, k) * (double)mat_at(Uf, k, j);
      CHECK(near(lu, mat_at(A, p->data[i], q->data[j]), 1e-5));
    }
  mat_free_perm(p);
  mat_free_perm(q);
  mat_free_mat(Lf);
  mat_free_mat(Uf);
  mat_free_mat(A);
  mat_free_mat(L);
  mat_free_mat(U);
  mat_free_mat(L1);
  mat_free_mat(x);
  mat_free_mat(b);
  mat_free_mat(b4);
  mat_free_mat(b1);
}

static void known_answers(void) {
  /* product: [1 2; 3 4] [5 6; 7 8] = [19 22; 43 50] */
  Mat *a = mat_from(2, 2, (mat_elem_t[]){1, 2, 3, 4});
  Mat *b = mat_from(2, 2, (mat_elem_t[]){5, 6, 7, 8});
  Mat *c = mat_mat(2, 2);
  mat_mul(c, a, b);
  CHECK(mat_at(c, 0, 0) == 19 && mat_at(c, 0, 1) == 22);
  CHECK(mat_at(c, 1, 0) == 43 && mat_at(c, 1, 1) == 50);
  Mat *r = mat_rmul(a, b);
  CHECK(mat_equals(r, c));
  mat_free_mat(r);
  /* C = 2 A B + 1 C */
  mat_gemm(c, 2, a, b, 1);
  CHECK(mat_at(c, 1, 1) == 150);
  mat_free_mat(a);
  mat_free_mat(b);
  mat_free_mat(c);

  /* solve: [2 3; 4 5] x = [8 14] -> x = [1 2] */
  Mat *A = mat_from(2, 2, (mat_elem_t[]){2, 3, 4, 5});
  Vec *rhs = mat_vec_from(2, (mat_elem_t[]){8, 14});
  Vec *x = mat_vec(2);
  mat_solve(x, A, rhs);
  CHECK(near(x->data[0], 1, 1e-5) && near(x->data[1], 2, 1e-5));
  CHECK(near(mat_det(A), -2, 1e-5));
  mat_free_mat(A);
  mat_free_mat(rhs);
  mat_free_mat(x);

  /* PLU of [1 2 3; 0 1 4; 5 6 0]: the first pivot is row 2 */
  Mat *M = mat_from(3, 3, (mat_elem_t[]){1, 2, 3, 0, 1, 4, 5, 6, 0});
  Mat *L = mat_mat(3, 3), *U = mat_mat(3, 3);
  Perm *p = mat_perm(3);
  int swaps = mat_plu(M, L, U, p);
  CHECK(p->data[0] == 2);
  CHECK(swaps >= 1);
  CHECK(mat_at(L, 0, 0) == 1 && mat_at(L, 0, 1) == 0 && mat_at(U, 1, 0) == 0);
  Mat *Minv = mat_mat(3, 3), *I3 = mat_mat(3, 3);
  mat_inv(Minv, M);
  mat_mul(I3, M, Minv);
  for (size_t i = 0; i < 3; ++i)
    for (size_t j = 0; j < 3; ++j) CHECK(near(mat_at(I3, i, j), i == j, 1e-4));
  mat_free_mat(Minv);
  mat_free_mat(I3);
  mat_free_perm(p);
  mat_free_mat(L);
  mat_free_mat(U);
  mat_free_mat(M);

  /* Cholesky: [4 2; 2 5] = L L^T with L = [2 0; 1 2]; [1 2; 2 1] is not PD */
  Mat *S = mat_from(2, 2, (mat_elem_t[]){4, 2, 2, 5});
  Mat *Ls = mat_mat(2, 2);
  CHECK(mat_chol(S, Ls) == 0);
  CHECK(near(mat_at(Ls, 0, 0), 2, 1e-6) && near(mat_at(Ls, 1, 0), 1, 1e-6));
  CHECK(near(mat_at(Ls, 1, 1), 2, 1e-6) && mat_at(Ls, 0, 1) == 0);
  Mat *N = mat_from(2, 2, (mat_elem_t[]){1, 2, 2, 1});
  CHECK(mat_chol(N, Ls) == -1);
  /* SPD solve: [4 2; 2 3] x = [2 1] -> x = [0.5 0] */
  Mat *S2 = mat_from(2, 2, (mat_elem_t[]){4, 2, 2, 3});
  Vec *b2 = mat_vec_from(2, (mat_elem_t[]){2, 1});
  Vec *x2 = mat_vec(2);
  CHECK(mat_solve_spd(x2, S2, b2) == 0);
  CHECK(near(x2->data[0], 0.5, 1e-5) && near(x2->data[1], 0, 1e-5));
  /* not PD: -1 and x untouched */
  x2->data[0] = 7;
  CHECK(mat_solve_spd(x2, N, b2) == -1 && x2->data[0] == 7);
  mat_free_mat(S);
  mat_free_mat(Ls);
  mat_free_mat(N);
  mat_free_mat(S2);
  mat_free_mat(b2);
  mat_free_mat(x2);
}

/* xorshift64*: a thread's own random stream */
static double uniform(unsigned long long *s) {
  *s ^= *s >> 12;
  *s ^= *s << 25;
  *s ^= *s >> 27;
  return (double)((*s * 2685821657736338717ull) >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

typedef struct {
  int id;
  int bad;
} Job;

static void *worker(void *arg) {
  Job *job = (Job *)arg;
  unsigned long long s = 0x9E3779B97F4A7C15ull * (unsigned long long)(job->id + 1);
  for (int it = 0; it < 4; ++it) {
    const size_t n = 300 + 97 * (size_t)job->id + 31 * (size_t)it;
    Mat *A = mat_mat(n, n);
    Vec *b = mat_vec(n), *x = mat_vec(n);
    for (size_t i = 0; i < n * n; ++i) A->data[i] = (mat_elem_t)uniform(&s);
    for (size_t i = 0; i < n; ++i) b->data[i] = (mat_elem_t)uniform(&s);
    mat_solve(x, A, b);
    /* normwise backward error ||Ax - b|| / (||A|| ||x|| + ||b||), inf norms */
    double rmax = 0, anorm = 0, xmax = 0, bmax = 0;
    for (size_t i = 0; i < n; ++i) {
      double acc = -(double)b->data[i], rs = 0;
      for (size_t j = 0; j < n; ++j) {
        acc += (double)mat_at(A, i, j) * (double)x->data[j];
        rs += fabs((double)mat_at(A, i, j));
      }
      if (fabs(acc) > rmax) rmax = fabs(acc);
      if (rs > anorm) anorm = rs;
      if (fabs((double)x->data[i]) > xmax) xmax = fabs((double)x->data[i]);
      if (fabs((double)b->data[i]) > bmax) bmax = fabs((double)b->data[i]);
    }
    const double eps = sizeof(mat_elem_t) == 4 ? 1.19e-7 : 2.22e-16;
    if (!(rmax / (anorm * xmax + bmax) <= 8.0 * (double)n * eps)) {
      fprintf(stderr, "thread %d: n=%zu backward error %g\n", job->id, n,
              rmax / (anorm * xmax + bmax));
      job->bad = 1;
    }
    mat_free_mat(A);
    mat_free_mat(b);
    mat_free_mat(x);
  }
  return NULL;
}

int main(int argc, char **argv) {
  const int threads = argc > 1 ? atoi(argv[1]) : 2;
  known_answers();
  triangular_and_full_pivot();
  CHECK(sizeof(linked) / sizeof(linked[0]) == 37);
  pthread_t tid[16];
  Job jobs[16];
  const int nt = threads < 16 ? threads : 16;
  for (int i = 0; i < nt; ++i) {
    jobs[i].id = i;
    jobs[i].bad = 0;
    pthread_create(&tid[i], NULL, worker, &jobs[i]);
  }
  for (int i = 0; i < nt; ++i) {
    pthread_join(tid[i], NULL);
    if (jobs[i].bad) ++failures;
  }
  printf("libmat_caller: %s (%d failures, %d threads)\n", failures ? "FAILED" : "ok", failures, nt);
  return failures ? 1 : 0;
}
