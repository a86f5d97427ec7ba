/* Runs libmat's API on every golden fixture (tests/golden/golden_*.lmfx,
 * read with tests/golden/lmfx.h) and compares with the recorded outputs.
 *
 *   exact: every output bit for bit (the fixtures were written by the
 *          reference, so this holds when the program runs on the reference
 *          itself: build with -DCHECK_REFERENCE and -I<libmat>);
 *   tol:   the GPU library's tolerances (MAT_HEADER "mat_gpu.h", linked to
 *          libmatgpu_f32.so / libmatgpu_f64.so): products within 2 K eps,
 *          PLU with PA = LU to 4 n eps and the reference's pivots and swap
 *          count where the pivots agree, full-pivot LU with identical p, q
 *          and swaps, Cholesky within 16 n eps and the same -1 cases,
 *          triangular solves within 64 n eps.
 *
 * Usage: fixture_check <file.lmfx> exact|tol.  Prints one line per failed
 * check and a summary; exit status 0 when all pass. */
#ifdef CHECK_REFERENCE
#define MATDEF static inline
#define MAT_IMPLEMENTATION
#include "mat.h"
#else
#include MAT_HEADER
#endif

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "lmfx.h"

typedef mat_elem_t T;
static const double EPS = sizeof(T) == 4 ? 1.1920929e-07 : 2.220446049250313e-16;
static int exact = 0, checks = 0, failures = 0;
static lmfx_file F;

#define FAIL(...)                         \
  do {                                    \
    fprintf(stderr, __VA_ARGS__);         \
    fputc('\n', stderr);                  \
    ++failures;                           \
  } while (0)

static lmfx_entry get(const char *tag, const char *suffix) {
  char name[128];
  snprintf(name, sizeof name, "%s_%s", tag, suffix);
  lmfx_entry e;
  if (lmfx_find(&F, name, &e) != 0) {
    FAIL("%s: missing entry", name);
    memset(&e, 0, sizeof e);
  }
  return e;
}

/* A fresh Mat holding a copy of a fixture matrix (or vector). */
static Mat *mat_of(const char *tag, const char *suffix) {
  lmfx_entry e = get(tag, suffix);
  Mat *m = mat_mat(e.rows ? e.rows : 1, e.cols ? e.cols : 1);
  if (e.data) memcpy(m->data, e.data, e.rows * e.cols * sizeof(T));
  return m;
}

static int64_t int_of(const char *tag, const char *suffix) {
  lmfx_entry e = get(tag, suffix);
  int64_t v = 0;
  if (e.data) memcpy(&v, e.data, 8);
  return v;
}

/* ||a - b||_F / ||b||_F (0 when both are zero) */
static double rel(const T *a, const T *b, size_t n) {
  double d = 0, r = 0;
  for (size_t i = 0; i < n; ++i) {
    d += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
    r += (double)b[i] * b[i];
  }
  return r > 0 ? sqrt(d / r) : sqrt(d);
}

/* outputs compared: bit for bit in exact mode, else within tol */
static void cmp(const char *what, const char *tag, const Mat *got, const Mat *want, double tol) {
  ++checks;
  const size_t n = got->rows * got->cols;
  if (exact) {
    if (memcmp(got->data, want->data, n * sizeof(T)) != 0) FAIL("%s %s: not bit-exact", tag, what);
  } else if (!(rel(got->data, want->data, n) <= tol)) {
    FAIL("%s %s: relative error %.3g > %.3g", tag, what, rel(got->data, want->data, n), tol);
  }
}

static int same_perm(const Perm *p, const char *tag, const char *suffix) {
  lmfx_entry e = get(tag, suffix);
  if (!e.data || e.rows != p->size) return 0;
  return memcmp(p->data, e.data, p->size * sizeof(size_t)) == 0;
}

static void check_gemm(const char *tag) {
  Mat *A = mat_of(tag, "A"), *B = mat_of(tag, "B"), *C = mat_of(tag, "C0"), *W = mat_of(tag, "C");
  mat_gemm(C, (T)0.7, A, B, (T)0.3);
  cmp("C", tag, C, W, 2.0 * A->cols * EPS);
  mat_free_mat(A), mat_free_mat(B), mat_free_mat(C), mat_free_mat(W);
}

static void check_plu(const char *tag, int with_solve) {
  Mat *A = mat_of(tag, "A"), *Lw = mat_of(tag, "L"), *Uw = mat_of(tag, "U");
  const size_t n = A->rows;
  Mat *L = mat_mat(n, n), *U = mat_mat(n, n);
  Perm *p = mat_perm(n);
  const int swaps = mat_plu(A, L, U, p);
  const int same = same_perm(p, tag, "p");
  ++checks;
  if (exact || same) {
    if (!same) FAIL("%s: pivots differ", tag);
    if (swaps != int_of(tag, "swaps")) FAIL("%s: swaps %d != %lld", tag, swaps,
                                            (long long)int_of(tag, "swaps"));
    cmp("L", tag, L, Lw, 64.0 * n * EPS);
    cmp("U", tag, U, Uw, 64.0 * n * EPS);
  } else {
    /* a near-tie chose another row: PA = LU with this library's p */
    double d = 0, r = 0;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        double lu = 0;
        for (size_t k = 0; k < n; ++k) lu += (double)mat_at(L, i, k) * mat_at(U, k, j);
        const double a = mat_at(A, p->data[i], j);
        d += (lu - a) * (lu - a);
        r += a * a;
      }
    if (!(sqrt(d / r) <= 4.0 * n * EPS)) FAIL("%s: ||PA - LU|| / ||A|| = %.3g", tag, sqrt(d / r));
  }
  if (with_solve) {
    Mat *b = mat_of(tag, "b"), *xw = mat_of(tag, "x");
    Vec *x = mat_vec(n);
    mat_solve(x, A, b);
    cmp("x", tag, x, xw, 1e3 * n * EPS);
    mat_free_mat(b), mat_free_mat(xw), mat_free_mat(x);
  }
  mat_free_mat(A), mat_free_mat(Lw), mat_free_mat(Uw), mat_free_mat(L), mat_free_mat(U);
  mat_free_perm(p);
}

static void check_lu(const char *tag) {
  Mat *A = mat_of(tag, "A"), *Lw = mat_of(tag, "L"), *Uw = mat_of(tag, "U");
  const size_t n = A->rows;
  Mat *L = mat_mat(n, n), *U = mat_mat(n, n);
  Perm *p = mat_perm(n), *q = mat_perm(n);
  const int swaps = mat_lu(A, L, U, p, q);
  ++checks;
  if (!same_perm(p, tag, "p") || !same_perm(q, tag, "q")) FAIL("%s: p or q differs", tag);
  if (swaps != int_of(tag, "swaps")) FAIL("%s: swaps differ", tag);
  cmp("L", tag, L, Lw, 32.0 * n * EPS);
  cmp("U", tag, U, Uw, 32.0 * n * EPS);
  mat_free_mat(A), mat_free_mat(Lw), mat_free_mat(Uw), mat_free_mat(L), mat_free_mat(U);
  mat_free_perm(p), mat_free_perm(q);
}

static void check_chol(const char *tag) {
  Mat *A = mat_of(tag, "A"), *Lw = mat_of(tag, "L"), *b = mat_of(tag, "b"), *xw = mat_of(tag, "x");
  const size_t n = A->rows;
  Mat *L = mat_mat(n, n);
  Vec *x = mat_vec(n);
  ++checks;
  if (mat_chol(A, L) != int_of(tag, "rc")) FAIL("%s: mat_chol return differs", tag);
  if (mat_solve_spd(x, A, b) != int_of(tag, "rs")) FAIL("%s: mat_solve_spd return differs", tag);
  cmp("L", tag, L, Lw, 16.0 * n * EPS);
  cmp("x", tag, x, xw, 1e3 * n * EPS);
  mat_free_mat(A), mat_free_mat(Lw), mat_free_mat(b), mat_free_mat(xw), mat_free_mat(L);
  mat_free_mat(x);
}

static void check_not_pd(const char *tag) {
  Mat *A = mat_of(tag, "A");
  const size_t n = A->rows;
  Mat *L = mat_mat(n, n);
  Vec *b = mat_vec(n), *x = mat_vec(n);
  for (size_t i = 0; i < n; ++i) b->data[i] = 1;
  ++checks;
  if (mat_chol(A, L) != int_of(tag, "rc")) FAIL("%s: mat_chol return differs", tag);
  if (mat_solve_spd(x, A, b) != int_of(tag, "rs")) FAIL("%s: mat_solve_spd return differs", tag);
  for (size_t i = 0; i < n; ++i)
    if (x->data[i] != 0) FAIL("%s: x written although A is not positive definite", tag);
  mat_free_mat(A), mat_free_mat(L), mat_free_mat(b), mat_free_mat(x);
}

static void check_tri(const char *tag) {
  Mat *Tm = mat_of(tag, "T"), *b = mat_of(tag, "b");
  const size_t n = Tm->rows;
  Mat *Lt = mat_mat(n, n), *Ut = mat_mat(n, n);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) {
      if (i >= j) mat_set_at(Lt, i, j, mat_at(Tm, i, j));
      if (i <= j) mat_set_at(Ut, i, j, mat_at(Tm, i, j));
    }
  Vec *x = mat_vec(n);
  const double tol = 64.0 * n * EPS;
  struct {
    const char *name;
    void (*fn)(Vec *, const Mat *, const Vec *);
    const Mat *m;
  } cases[] = {{"tril", mat_solve_tril, Lt}, {"triu", mat_solve_triu, Ut}, {"trilt", mat_solve_trilt, Lt},
               {"tril_unit", mat_solve_tril_unit, Lt}};
  for (int c = 0; c < 4; ++c) {
    /* random unit-lower systems are exponentially ill-conditioned: their
       solutions are compared bit for bit only */
    if (!exact && c == 3) continue;
    cases[c].fn(x, cases[c].m, b);
    Mat *w = mat_of(tag, cases[c].name);
    cmp(cases[c].name, tag, x, w, tol);
    mat_free_mat(w);
  }
  mat_free_mat(Tm), mat_free_mat(b), mat_free_mat(Lt), mat_free_mat(Ut), mat_free_mat(x);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <file.lmfx> exact|tol\n", argv[0]);
    return 2;
  }
  exact = strcmp(argv[2], "exact") == 0;
  if (lmfx_open(&F, argv[1]) != 0) {
    fprintf(stderr, "%s: cannot read\n", argv[1]);
    return 2;
  }
  lmfx_entry e;
  for (uint32_t i = 0; i < F.count && lmfx_entry_at(&F, i, &e) == 0; ++i) {
    char name[128];
    const size_t len = e.name_len < sizeof name - 1 ? e.name_len : sizeof name - 1;
    memcpy(name, e.name, len);
    name[len] = 0;
    char *us = strrchr(name, '_');
    if (!us) continue;
    const char *suffix = us + 1;
    *us = 0; /* name is now the fixture's tag */
    if (strncmp(name, "gemm_", 5) == 0 && strcmp(suffix, "A") == 0) check_gemm(name);
    else if ((strncmp(name, "plu_", 4) == 0 || strcmp(name, "cfg1") == 0) && strcmp(suffix, "A") == 0)
      check_plu(name, 1);
    else if ((strcmp(name, "zerocol30") == 0 || strcmp(name, "rank2_40") == 0 ||
              strcmp(name, "sing3") == 0) && strcmp(suffix, "A") == 0 && exact)
      check_plu(name, 0);
    else if (strncmp(name, "lu_", 3) == 0 && strcmp(suffix, "A") == 0) check_lu(name);
    else if (strncmp(name, "chol_", 5) == 0 && strcmp(suffix, "A") == 0) check_chol(name);
    else if ((strcmp(name, "notpd2") == 0 || strcmp(name, "indef40") == 0) && strcmp(suffix, "A") == 0)
      check_not_pd(name);
    else if (strncmp(name, "tri_", 4) == 0 && strcmp(suffix, "T") == 0) check_tri(name);
  }
  lmfx_close(&F);
  printf("fixture_check %s: %d checks, %d failures\n", exact ? "exact" : "tol", checks, failures);
  return failures ? 1 : 0;
}
