/*
 * Thin exporter around the reference library, compiled from the
 * reference's own single header (mat.h) where it lies.  TEST
 * INFRASTRUCTURE ONLY: the output goes to oracle/_ref/ and is used to pin
 * the oracle restatement (and, optionally, as the CPU baseline).
 *
 * mat.h is included the way the reference's own unit tests include it
 * (MATDEF static inline + MAT_IMPLEMENTATION, tests/test_mat_plu.c:1-3);
 * with external linkage the x86 link fails on an undefined
 * mat__norm_fro_fast_scalar, so every entry point here is a wrapper with
 * the same raw-pointer signature as the oracle's.  Built without -march
 * so MAT_DISPATCH picks the scalar kernels (mat.h:899-906).
 */
#define MATDEF static inline
#define MAT_IMPLEMENTATION
#include "mat.h"

#ifdef MAT_DOUBLE_PRECISION
#define REF(name) ref_d_##name
#else
#define REF(name) ref_s_##name
#endif

typedef mat_elem_t elem;

static Mat wrap(size_t r, size_t c, const elem *p) {
  Mat m;
  m.rows = r;
  m.cols = c;
  m.data = (elem *)p;
  return m;
}

void REF(gemm)(size_t M, size_t K, size_t N, elem alpha, const elem *A,
               const elem *B, elem beta, elem *C) {
  Mat a = wrap(M, K, A), b = wrap(K, N, B), c = wrap(M, N, C);
  mat_gemm(&c, alpha, &a, &b, beta);
}

void REF(syrk)(size_t n, size_t k, elem alpha, const elem *A, elem beta,
               elem *C, int upper, int trans) {
  Mat a = trans ? wrap(k, n, A) : wrap(n, k, A), c = wrap(n, n, C);
  if (trans) mat_syrk_t(&c, &a, alpha, beta, upper ? 'U' : 'L');
  else mat_syrk(&c, &a, alpha, beta, upper ? 'U' : 'L');
}

void REF(mul)(size_t M, size_t K, size_t N, const elem *A, const elem *B,
              elem *C) {
  Mat a = wrap(M, K, A), b = wrap(K, N, B), c = wrap(M, N, C);
  mat_mul(&c, &a, &b);
}

int REF(plu)(size_t n, const elem *A, elem *L, elem *U, size_t *p) {
  Mat a = wrap(n, n, A), l = wrap(n, n, L), u = wrap(n, n, U);
  Perm pp;
  pp.data = p;
  pp.size = n;
  return mat_plu(&a, &l, &u, &pp);
}

int REF(lu)(size_t n, const elem *A, elem *L, elem *U, size_t *p,
            size_t *q) {
  Mat a = wrap(n, n, A), l = wrap(n, n, L), u = wrap(n, n, U);
  Perm pp, qq;
  pp.data = p;
  pp.size = n;
  qq.data = q;
  qq.size = n;
  return mat_lu(&a, &l, &u, &pp, &qq);
}

int REF(chol)(size_t n, const elem *A, elem *L) {
  Mat a = wrap(n, n, A), l = wrap(n, n, L);
  return mat_chol(&a, &l);
}

#define TRI(fn)                                                               \
  void REF(fn)(size_t n, const elem *T, const elem *b, elem *x) {             \
    Mat t = wrap(n, n, T), bb = wrap(n, 1, b), xx = wrap(n, 1, x);            \
    mat_##fn(&xx, &t, &bb);                                                   \
  }
TRI(solve_tril)
TRI(solve_tril_unit)
TRI(solve_triu)
TRI(solve_trilt)

void REF(solve)(size_t n, const elem *A, const elem *b, elem *x) {
  Mat a = wrap(n, n, A), bb = wrap(n, 1, b), xx = wrap(n, 1, x);
  mat_solve(&xx, &a, &bb);
}

int REF(solve_spd)(size_t n, const elem *A, const elem *b, elem *x) {
  Mat a = wrap(n, n, A), bb = wrap(n, 1, b), xx = wrap(n, 1, x);
  return mat_solve_spd(&xx, &a, &bb);
}

/* Factor once with mat_plu, then per right-hand side permute and run
 * mat_solve_tril_unit + mat_solve_triu: mat_solve's own internals
 * (mat.h:4823-4830) without refactorizing for every column. */
void REF(solve_multi)(size_t n, size_t nrhs, const elem *A, const elem *B,
                      elem *X) {
  Mat *L = mat_mat(n, n), *U = mat_mat(n, n);
  Perm *p = mat_perm(n);
  Vec *pb = mat_vec(n), *z = mat_vec(n);
  Mat a = wrap(n, n, A);
  mat_plu(&a, L, U, p);
  for (size_t c = 0; c < nrhs; c++) {
    for (size_t i = 0; i < n; i++) pb->data[i] = B[c * n + p->data[i]];
    Mat xx = wrap(n, 1, &X[c * n]);
    mat_solve_tril_unit(z, L, pb);
    mat_solve_triu(&xx, U, z);
  }
  mat_free_mat(L);
  mat_free_mat(U);
  mat_free_perm(p);
  mat_free_mat(pb);
  mat_free_mat(z);
}

/* Cholesky once, then per right-hand side mat_solve_tril +
 * mat_solve_trilt: mat_solve_spd's internals (mat.h:4854-4865). */
int REF(chol_solve_multi)(size_t n, size_t nrhs, const elem *A, const elem *B,
                          elem *X) {
  Mat *L = mat_mat(n, n);
  Vec *y = mat_vec(n);
  Mat a = wrap(n, n, A);
  int r = mat_chol(&a, L);
  if (r == 0) {
    for (size_t c = 0; c < nrhs; c++) {
      Mat bb = wrap(n, 1, &B[c * n]), xx = wrap(n, 1, &X[c * n]);
      mat_solve_tril(y, L, &bb);
      mat_solve_trilt(&xx, L, y);
    }
  }
  mat_free_mat(L);
  mat_free_mat(y);
  return r;
}

/* One mat_solve (kind 0) or mat_solve_spd (kind 1) per matrix, in order. */
void REF(batch_solve)(size_t count, const int *n, const signed char *kind,
                      const elem *A, const long long *a_off, const elem *b,
                      elem *x, const long long *v_off, int *info) {
  for (size_t i = 0; i < count; i++) {
    const size_t ni = (size_t)n[i];
    Mat a = wrap(ni, ni, A + a_off[i]), bb = wrap(ni, 1, b + v_off[i]),
        xx = wrap(ni, 1, x + v_off[i]);
    if (kind[i] == 0) {
      mat_solve(&xx, &a, &bb);
      info[i] = 0;
    } else {
      info[i] = mat_solve_spd(&xx, &a, &bb);
    }
  }
}

void REF(inv)(size_t n, const elem *A, elem *out) {
  Mat a = wrap(n, n, A), o = wrap(n, n, out);
  mat_inv(&o, &a);
}

elem REF(det)(size_t n, const elem *A) {
  Mat a = wrap(n, n, A);
  return mat_det(&a);
}

void REF(qr)(size_t m, size_t n, const elem *A, elem *Q, elem *R) {
  Mat a = wrap(m, n, A), q = wrap(m, m, Q), r = wrap(m, n, R);
  mat_qr(&a, &q, &r);
}

void REF(qr_r)(size_t m, size_t n, const elem *A, elem *R) {
  Mat a = wrap(m, n, A), r = wrap(m, n, R);
  mat_qr_r(&a, &r);
}
