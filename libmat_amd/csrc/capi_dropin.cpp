// Drop-in libmat entry points (include/mat_gpu.h, part 1).  Compiled once
// per element type: MAT_DOUBLE_PRECISION selects double, as in mat.h.
//
// Each call uploads the host Mats, runs the HIP path on the library's
// stream, downloads the outputs and returns synchronously.  Contract
// violations the reference checks with MAT_ASSERT (mat.h:327-357) and HIP
// failures abort with a message; nothing falls back to the CPU.
#include <cmath>
#include <cstring>
#include <vector>

#include "mat_gpu.h"
#include "ops.hpp"

using mg::View;
using T = mat_elem_t;

namespace {

constexpr bool kF32 = sizeof(T) == 4;
const T kEps = kF32 ? (T)1e-6f : (T)1e-9;  // MAT_DEFAULT_EPSILON, mat.h:168,176
constexpr int kPluLeaf = 24;                // MAT_PLU_BLOCK_SIZE, mat.h:4196

[[noreturn]] void fail(const char *fn, const std::string &msg) {
  fprintf(stderr, "libmat_amd: %s: %s\n", fn, msg.c_str());
  fflush(stderr);
  abort();
}

#define REQUIRE(cond)                                                \
  do {                                                               \
    if (!(cond)) fail(__func__, "assertion failed: " #cond);         \
  } while (0)
#define REQUIRE_MAT(m)                                               \
  do {                                                               \
    REQUIRE((m) != NULL);                                            \
    REQUIRE((m)->data != NULL);                                      \
    REQUIRE((m)->rows > 0 && (m)->cols > 0);                         \
  } while (0)
#define REQUIRE_SQUARE(m)                                            \
  do {                                                               \
    REQUIRE_MAT(m);                                                  \
    REQUIRE((m)->rows == (m)->cols);                                 \
  } while (0)

template <typename F>
auto run(const char *fn, F &&f) -> decltype(f()) {
  try {
    return f();
  } catch (const mg::Error &e) {
    fail(fn, e.msg);
  }
}

size_t elems(const Mat *m) { return m->rows * m->cols; }

T *upload(hipStream_t s, const char *slot, const Mat *m) {
  const size_t bytes = elems(m) * sizeof(T);
  T *d = (T *)mg::workspace(s).get(slot, bytes);
  MG_HIP(hipMemcpyAsync(d, m->data, bytes, hipMemcpyHostToDevice, s));
  return d;
}

void download(hipStream_t s, Mat *m, const T *d) {
  MG_HIP(hipMemcpyAsync(m->data, d, elems(m) * sizeof(T), hipMemcpyDeviceToHost, s));
}

void finish(hipStream_t s) { mg::sync_and_check(s); }

// PLU of host A into the device row-major working matrix; returns it and
// fills the host ipiv.
T *plu_device(hipStream_t s, const Mat *A, int **perm_dev, std::vector<int> &ipiv_h) {
  const int n = (int)A->rows;
  mg::Workspace &ws = mg::workspace(s);
  T *dA = upload(s, "dropin.A", A);
  T *W = (T *)ws.get("dropin.W", (size_t)n * n * sizeof(T));
  int *ipiv = (int *)ws.get("dropin.ipiv", (size_t)n * 4);
  int *perm = (int *)ws.get("dropin.perm", (size_t)n * 4);
  mg::transpose<T>(s, n, n, dA, n, W, n);
  mg::LuOpts o;
  o.leaf = kPluLeaf;
  mg::getrf_rowmajor_perm<T>(s, n, W, n, ipiv, perm, kEps, o);
  ipiv_h.resize(n);
  MG_HIP(hipMemcpyAsync(ipiv_h.data(), ipiv, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  if (perm_dev) *perm_dev = perm;
  return W;
}

}  // namespace

namespace mg {
template <typename U>
int tri_block_size();
template <typename U>
void tri_invert_blocks(hipStream_t, bool, bool, int, View<const U>, U *);
template <typename U>
void tri_solve_blocked(hipStream_t, bool, int, int, View<const U>, const U *, U *, long, const U *);
template <typename U>
void getrs_rowmajor(hipStream_t, int, int, const U *, long, const int *, const U *, long, U *, long);
template <typename U>
void potrs_lower(hipStream_t, int, int, const U *, long, U *, long);
}  // namespace mg

extern "C" {

int mat_gpu_elem_size(void) { return (int)sizeof(T); }

// ------------------------------------------------------------ products --
void mat_gemm(Mat *C, T alpha, const Mat *A, const Mat *B, T beta) {
  REQUIRE_MAT(C);
  REQUIRE_MAT(A);
  REQUIRE_MAT(B);
  REQUIRE(A->cols == B->rows);
  REQUIRE(C->rows == A->rows && C->cols == B->cols);
  run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const long M = (long)A->rows, K = (long)A->cols, N = (long)B->cols;
    T *dA = upload(s, "dropin.A", A);
    T *dB = upload(s, "dropin.B", B);
    T *dC = beta != T(0) ? upload(s, "dropin.C", C)
                         : (T *)mg::workspace(s).get("dropin.C", elems(C) * sizeof(T));
    mg::gemm<T>(s, M, N, K, alpha, View<const T>{dA, 1, M}, View<const T>{dB, 1, K}, beta,
                View<T>{dC, 1, M}, false);
    download(s, C, dC);
    finish(s);
  });
}

void mat_mul(Mat *out, const Mat *a, const Mat *b) {
  REQUIRE_MAT(out);
  REQUIRE_MAT(a);
  REQUIRE_MAT(b);
  REQUIRE(a->cols == b->rows);
  mat_gemm(out, 1, a, b, 0);
}

Mat *mat_rmul(const Mat *a, const Mat *b) {
  REQUIRE_MAT(a);
  REQUIRE_MAT(b);
  REQUIRE(a->cols == b->rows);
  Mat *r = mat_mat(a->rows, b->cols);
  mat_mul(r, a, b);
  return r;
}

// C = alpha*op(A)*op(A)^T + beta*C on one triangle of the square C
// (mat_syrk: op(A) = A, mat.h:5670-5677; mat_syrk_t: op(A) = A^T,
// mat.h:5727-5734; semantics of their scalar kernels, mat.h:5641-5668 and
// 5698-5725): uplo 'L' or 'l' updates row >= col, any other value row <=
// col; the other triangle is left as it is.
static void syrk_dropin(const char *fn, Mat *C, const Mat *A, T alpha, T beta, char uplo,
                        bool trans) {
  run(fn, [&] {
    hipStream_t s = mg::default_stream();
    const long n = (long)C->rows, k = (long)(trans ? A->rows : A->cols);
    T *dA = upload(s, "dropin.A", A);
    T *dC = upload(s, "dropin.C", C);
    // op(A) is n x k: A itself (column-major, ld n) or A^T read from A (k x n, ld k)
    const View<const T> a = trans ? View<const T>{dA, k, 1} : View<const T>{dA, 1, n};
    const int tri = (uplo == 'L' || uplo == 'l') ? 1 : 2;
    mg::gemm<T>(s, n, n, k, alpha, a, a.t(), beta, View<T>{dC, 1, n}, tri);
    download(s, C, dC);
    finish(s);
  });
}

void mat_syrk(Mat *C, const Mat *A, T alpha, T beta, char uplo) {
  REQUIRE_MAT(C);
  REQUIRE_MAT(A);
  REQUIRE(C->rows == C->cols);
  REQUIRE(C->rows == A->rows);
  syrk_dropin(__func__, C, A, alpha, beta, uplo, false);
}

void mat_syrk_t(Mat *C, const Mat *A, T alpha, T beta, char uplo) {
  REQUIRE_MAT(C);
  REQUIRE_MAT(A);
  REQUIRE(C->rows == C->cols);
  REQUIRE(C->rows == A->cols);
  syrk_dropin(__func__, C, A, alpha, beta, uplo, true);
}

// -------------------------------------------------------------------- QR --
// Householder QR (mat_qr, mat.h:3907-4063; mat_qr_r, mat.h:4066-4192): the
// reference asserts m >= n and the output shapes (mat.h:3914-3916,
// 4072-4073).
void mat_qr(const Mat *A, Mat *Q, Mat *R) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(Q);
  REQUIRE_MAT(R);
  REQUIRE(A->rows >= A->cols);
  REQUIRE(Q->rows == A->rows && Q->cols == A->rows);
  REQUIRE(R->rows == A->rows && R->cols == A->cols);
  run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const int m = (int)A->rows, n = (int)A->cols;
    T *dR = upload(s, "dropin.C", A);
    T *dQ = (T *)mg::workspace(s).get("dropin.B", (size_t)m * m * sizeof(T));
    mg::geqrf<T>(s, m, n, dR, m, dQ, kEps);
    download(s, R, dR);
    download(s, Q, dQ);
    finish(s);
  });
}

void mat_qr_r(const Mat *A, Mat *R) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(R);
  REQUIRE(A->rows >= A->cols);
  REQUIRE(R->rows == A->rows && R->cols == A->cols);
  run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const int m = (int)A->rows, n = (int)A->cols;
    T *dR = upload(s, "dropin.C", A);
    mg::geqrf<T>(s, m, n, dR, m, (T *)nullptr, kEps);
    download(s, R, dR);
    finish(s);
  });
}

// -------------------------------------------------------- factorizations --
int mat_plu(const Mat *A, Mat *L, Mat *U, Perm *p) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(L);
  REQUIRE_MAT(U);
  REQUIRE(p != NULL);
  REQUIRE_SQUARE(A);
  const size_t n = A->rows;
  REQUIRE(L->rows == n && L->cols == n);
  REQUIRE(U->rows == n && U->cols == n);
  REQUIRE(p->size == n);
  return run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    std::vector<int> ipiv;
    T *W = plu_device(s, A, nullptr, ipiv);
    mg::Workspace &ws = mg::workspace(s);
    T *dL = (T *)ws.get("dropin.L", n * n * sizeof(T));
    T *dU = (T *)ws.get("dropin.U", n * n * sizeof(T));
    mg::split_lu_rowmajor<T>(s, (int)n, W, (long)n, dL, dU);
    download(s, L, dL);
    download(s, U, dU);
    finish(s);
    return mg::ipiv_to_perm((int)n, ipiv.data(), p->data);
  });
}

int mat_lu(const Mat *A, Mat *L, Mat *U, Perm *p, Perm *q) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(L);
  REQUIRE_MAT(U);
  REQUIRE(p != NULL && q != NULL);
  const size_t n = A->rows;
  REQUIRE(A->cols == n);
  REQUIRE(L->rows == n && L->cols == n);
  REQUIRE(U->rows == n && U->cols == n);
  REQUIRE(p->size == n && q->size == n);
  return run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    mg::Workspace &ws = mg::workspace(s);
    T *dA = upload(s, "dropin.A", A);
    int *rec = (int *)ws.get("dropin.rec", 3 * n * sizeof(int) + 4);
    mg::getrf_full<T>(s, (int)n, dA, (long)n, rec, kEps);
    T *dL = (T *)ws.get("dropin.L", n * n * sizeof(T));
    T *dU = (T *)ws.get("dropin.U", n * n * sizeof(T));
    mg::split_lu_colmajor<T>(s, (int)n, dA, (long)n, dL, dU);
    std::vector<int> h(3 * n + 1);
    if (n > 1) MG_HIP(hipMemcpyAsync(h.data(), rec, 3 * (n - 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    download(s, L, dL);
    download(s, U, dU);
    finish(s);
    for (size_t i = 0; i < n; ++i) p->data[i] = q->data[i] = i;
    int swaps = 0;
    for (size_t k = 0; k + 1 < n; ++k) {
      if (h[3 * k + 2]) continue;
      const size_t pr = (size_t)h[3 * k], pc = (size_t)h[3 * k + 1];
      if (pr != k) {
        std::swap(p->data[k], p->data[pr]);
        ++swaps;
      }
      if (pc != k) {
        std::swap(q->data[k], q->data[pc]);
        ++swaps;
      }
    }
    return swaps;
  });
}

int mat_chol(const Mat *A, Mat *L) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(L);
  REQUIRE_SQUARE(A);
  const size_t n = A->rows;
  REQUIRE(L->rows == n && L->cols == n);
  return run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    mg::Workspace &ws = mg::workspace(s);
    T *dA = upload(s, "dropin.A", A);
    T *dL = (T *)ws.get("dropin.L", n * n * sizeof(T));
    int *info = (int *)ws.get("dropin.info", 4);
    mg::copy_lower<T>(s, (int)n, dA, (long)n, dL, (long)n);
    mg::potrf_lower<T>(s, (int)n, dL, (long)n, info);
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
    download(s, L, dL);
    finish(s);
    return h != 0 ? -1 : 0;
  });
}

// ---------------------------------------------------------------- solves --
void mat_solve(Vec *x, const Mat *A, const Vec *b) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(x);
  REQUIRE_MAT(b);
  REQUIRE_SQUARE(A);
  REQUIRE(b->rows == A->rows && "b must have same rows as A");
  REQUIRE(x->rows == A->cols && "x must have same rows as A cols");
  run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const int n = (int)A->rows;
    int *perm = nullptr;
    std::vector<int> ipiv;
    T *W = plu_device(s, A, &perm, ipiv);
    mg::Workspace &ws = mg::workspace(s);
    T *db = (T *)ws.get("dropin.b", (size_t)n * sizeof(T));
    T *dx = (T *)ws.get("dropin.x", (size_t)n * sizeof(T));
    MG_HIP(hipMemcpyAsync(db, b->data, (size_t)n * sizeof(T), hipMemcpyHostToDevice, s));
    mg::getrs_rowmajor<T>(s, n, 1, W, n, perm, db, n, dx, n);
    MG_HIP(hipMemcpyAsync(x->data, dx, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s));
    finish(s);
  });
}

int mat_solve_spd(Vec *x, const Mat *A, const Vec *b) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(x);
  REQUIRE_MAT(b);
  REQUIRE_SQUARE(A);
  REQUIRE(b->rows == A->rows && "b must have same rows as A");
  REQUIRE(x->rows == A->cols && "x must have same rows as A cols");
  return run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const int n = (int)A->rows;
    mg::Workspace &ws = mg::workspace(s);
    T *dA = upload(s, "dropin.A", A);
    T *dL = (T *)ws.get("dropin.L", (size_t)n * n * sizeof(T));
    int *info = (int *)ws.get("dropin.info", 4);
    mg::copy_lower<T>(s, n, dA, n, dL, n);
    mg::potrf_lower<T>(s, n, dL, n, info);
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
    finish(s);
    if (h != 0) return -1;  // not positive definite: x untouched (mat.h:4855-4859)
    T *dx = (T *)ws.get("dropin.x", (size_t)n * sizeof(T));
    MG_HIP(hipMemcpyAsync(dx, b->data, (size_t)n * sizeof(T), hipMemcpyHostToDevice, s));
    mg::potrs_lower<T>(s, n, 1, dL, n, dx, n);
    MG_HIP(hipMemcpyAsync(x->data, dx, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s));
    finish(s);
    return 0;
  });
}

static void tri_solve(const char *fn, Vec *x, const Mat *Tm, const Vec *b, bool lower, bool unit,
                      bool transpose) {
  run(fn, [&] {
    hipStream_t s = mg::default_stream();
    const int n = (int)Tm->rows;
    T *dT = upload(s, "dropin.A", Tm);
    T *dx = (T *)mg::workspace(s).get("dropin.x", (size_t)n * sizeof(T));
    MG_HIP(hipMemcpyAsync(dx, b->data, (size_t)n * sizeof(T), hipMemcpyHostToDevice, s));
    View<const T> v{dT, 1, n};
    if (transpose) v = v.t();
    const int tb = mg::tri_block_size<T>();
    T *D = (T *)mg::workspace(s).get("dropin.D", (size_t)((n + tb - 1) / tb) * tb * tb * sizeof(T));
    mg::tri_invert_blocks<T>(s, lower, unit, n, v, D);
    mg::tri_solve_blocked<T>(s, lower, n, 1, v, D, dx, n, nullptr);
    MG_HIP(hipMemcpyAsync(x->data, dx, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s));
    finish(s);
  });
}

#define TRI_CHECKS(Tm)                 \
  REQUIRE_MAT(Tm);                     \
  REQUIRE_MAT(x);                      \
  REQUIRE_MAT(b);                      \
  REQUIRE_SQUARE(Tm);                  \
  REQUIRE(b->rows == (Tm)->rows);      \
  REQUIRE(x->rows == (Tm)->rows)

void mat_solve_tril(Vec *x, const Mat *L, const Vec *b) {
  TRI_CHECKS(L);
  tri_solve(__func__, x, L, b, true, false, false);
}
void mat_solve_tril_unit(Vec *x, const Mat *L, const Vec *b) {
  TRI_CHECKS(L);
  tri_solve(__func__, x, L, b, true, true, false);
}
void mat_solve_triu(Vec *x, const Mat *U, const Vec *b) {
  TRI_CHECKS(U);
  tri_solve(__func__, x, U, b, false, false, false);
}
void mat_solve_trilt(Vec *x, const Mat *L, const Vec *b) {
  TRI_CHECKS(L);
  tri_solve(__func__, x, L, b, false, false, true);
}

// ------------------------------------------------------ PLU consumers ----
T mat_det(const Mat *A) {
  REQUIRE_MAT(A);
  REQUIRE_SQUARE(A);
  return run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    const int n = (int)A->rows;
    std::vector<int> ipiv;
    T *W = plu_device(s, A, nullptr, ipiv);
    std::vector<T> diag(n);
    MG_HIP(hipMemcpy2DAsync(diag.data(), sizeof(T), W, (size_t)(n + 1) * sizeof(T), sizeof(T), n,
                            hipMemcpyDeviceToHost, s));
    finish(s);
    std::vector<size_t> p(n);
    const int swaps = mg::ipiv_to_perm(n, ipiv.data(), p.data());
    T det = (swaps % 2 == 0) ? T(1) : T(-1);  // mat.h:6241-6244
    for (int i = 0; i < n; ++i) det *= diag[i];
    return det;
  });
}

void mat_inv(Mat *out, const Mat *A) {
  REQUIRE_MAT(A);
  REQUIRE_MAT(out);
  REQUIRE_SQUARE(A);
  const size_t n = A->rows;
  REQUIRE(out->rows == n && out->cols == n);
  run(__func__, [&] {
    hipStream_t s = mg::default_stream();
    int *perm = nullptr;
    std::vector<int> ipiv;
    T *W = plu_device(s, A, &perm, ipiv);
    mg::Workspace &ws = mg::workspace(s);
    std::vector<T> eye(n * n, T(0));
    for (size_t i = 0; i < n; ++i) eye[i * n + i] = T(1);
    T *dI = (T *)ws.get("dropin.B", n * n * sizeof(T));
    T *dX = (T *)ws.get("dropin.C", n * n * sizeof(T));
    MG_HIP(hipMemcpyAsync(dI, eye.data(), n * n * sizeof(T), hipMemcpyHostToDevice, s));
    mg::getrs_rowmajor<T>(s, (int)n, (int)n, W, (long)n, perm, dI, (long)n, dX, (long)n);
    download(s, out, dX);
    finish(s);
  });
}

// The host helpers (mat_empty ... mat_equals_tol) are in mat_host.cpp.

}  // extern "C"
