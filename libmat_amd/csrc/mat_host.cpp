// libmat's constructors and accessors that a caller uses around the hot
// path (mat.h:398-480), pure host code with no HIP dependency, so that it
// can also be built and checked on its own (the sanitizer test builds it
// with AddressSanitizer and UBSan).  Compiled once per element type, like
// capi_dropin.cpp: MAT_DOUBLE_PRECISION selects double, as in mat.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mat_gpu.h"

using T = mat_elem_t;

namespace {

constexpr bool kF32 = sizeof(T) == 4;
const T kEps = kF32 ? (T)1e-6f : (T)1e-9;  // MAT_DEFAULT_EPSILON, mat.h:168,176

[[noreturn]] void fail(const char *fn, const char *msg) {
  fprintf(stderr, "libmat_amd: %s: %s\n", fn, msg);
  fflush(stderr);
  abort();
}

#define REQUIRE(cond)                                        \
  do {                                                       \
    if (!(cond)) fail(__func__, "assertion failed: " #cond); \
  } while (0)
#define REQUIRE_MAT(m)                               \
  do {                                               \
    REQUIRE((m) != NULL);                            \
    REQUIRE((m)->data != NULL);                      \
    REQUIRE((m)->rows > 0 && (m)->cols > 0);         \
  } while (0)
#define REQUIRE_SQUARE(m)              \
  do {                                 \
    REQUIRE_MAT(m);                    \
    REQUIRE((m)->rows == (m)->cols);   \
  } while (0)

size_t elems(const Mat *m) { return m->rows * m->cols; }

}  // namespace

extern "C" {

// Same semantics as mat.h:1266-1422 (calloc'd storage, row-major inits).
Mat *mat_empty(size_t rows, size_t cols) {
  REQUIRE(rows > 0);
  REQUIRE(cols > 0);
  Mat *m = (Mat *)malloc(sizeof(Mat));
  m->rows = rows;
  m->cols = cols;
  m->data = NULL;
  return m;
}
Mat *mat_mat(size_t rows, size_t cols) {
  Mat *m = mat_empty(rows, cols);
  m->data = (T *)calloc(rows * cols, sizeof(T));
  return m;
}
void mat_init(Mat *out, const T *values) {
  REQUIRE_MAT(out);
  REQUIRE(values != NULL);
  for (size_t i = 0; i < out->rows; ++i)
    for (size_t j = 0; j < out->cols; ++j) out->data[j * out->rows + i] = values[i * out->cols + j];
}
Mat *mat_from(size_t rows, size_t cols, const T *values) {
  Mat *m = mat_mat(rows, cols);
  mat_init(m, values);
  return m;
}
void mat_free_mat(Mat *m) {
  REQUIRE_MAT(m);
  free(m->data);
  free(m);
}
Mat *mat_zeros(size_t rows, size_t cols) { return mat_mat(rows, cols); }
void mat_eye(Mat *out) {
  REQUIRE_SQUARE(out);
  memset(out->data, 0, elems(out) * sizeof(T));
  for (size_t i = 0; i < out->rows; ++i) out->data[i * out->rows + i] = 1;
}
Mat *mat_reye(size_t dim) {
  REQUIRE(dim > 0);
  Mat *m = mat_mat(dim, dim);
  for (size_t i = 0; i < dim; ++i) m->data[i * dim + i] = 1;
  return m;
}
Vec *mat_vec(size_t dim) { return mat_mat(dim, 1); }
Vec *mat_vec_from(size_t dim, const T *values) { return mat_from(dim, 1, values); }
Perm *mat_perm(size_t n) {
  Perm *p = (Perm *)malloc(sizeof(Perm));
  p->data = (size_t *)malloc(n * sizeof(size_t));
  p->size = n;
  return p;
}
void mat_free_perm(Perm *p) {
  free(p->data);
  free(p);
}
void mat_perm_identity(Perm *p) {
  for (size_t i = 0; i < p->size; ++i) p->data[i] = i;
}
Mat *mat_perm_mat(const Perm *p) {
  const size_t n = p->size;
  Mat *m = mat_zeros(n, n);
  for (size_t i = 0; i < n; ++i) m->data[p->data[i] * n + i] = 1;
  return m;
}
void mat_deep_copy(Mat *out, const Mat *src) {
  REQUIRE_MAT(out);
  REQUIRE_MAT(src);
  REQUIRE(out->rows == src->rows && out->cols == src->cols);
  memcpy(out->data, src->data, elems(src) * sizeof(T));
}
Mat *mat_rdeep_copy(const Mat *m) {
  REQUIRE_MAT(m);
  Mat *r = mat_mat(m->rows, m->cols);
  memcpy(r->data, m->data, elems(m) * sizeof(T));
  return r;
}
T mat_at(const Mat *m, size_t row, size_t col) {
  REQUIRE_MAT(m);
  return m->data[col * m->rows + row];
}
void mat_set_at(Mat *m, size_t row, size_t col, T value) {
  REQUIRE_MAT(m);
  m->data[col * m->rows + row] = value;
}
bool mat_equals_tol(const Mat *a, const Mat *b, T eps) {
  REQUIRE_MAT(a);
  REQUIRE_MAT(b);
  if (a->rows != b->rows || a->cols != b->cols) return false;
  for (size_t i = 0; i < elems(a); ++i) {
    T d = a->data[i] - b->data[i];
    if (d < 0) d = -d;
    if (d > eps) return false;
  }
  return true;
}
bool mat_equals(const Mat *a, const Mat *b) { return mat_equals_tol(a, b, kEps); }

}  // extern "C"
