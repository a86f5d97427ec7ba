// Device-resident extension API (include/mat_gpu.h, part 2).  Compiled once
// and linked into both flavour libraries.
#include <vector>

#include "mat_gpu.h"
#include "ops.hpp"

namespace mg {
int gesv_f64(hipStream_t, int, int, const double *, long, const double *, long, double *, long, int);
int gesv_mixed(hipStream_t, int, int, const double *, long, const double *, long, double *, long,
               const mat_gpu_ir_opts *, mat_gpu_ir_status *);
int posv_f64(hipStream_t, int, int, const double *, long, const double *, long, double *, long);
int posv_mixed(hipStream_t, int, int, const double *, long, const double *, long, double *, long,
               const mat_gpu_ir_opts *, mat_gpu_ir_status *);
template <typename T>
void getrs_rowmajor(hipStream_t, int, int, const T *, long, const int *, const T *, long, T *, long);
template <typename T>
void potrs_lower(hipStream_t, int, int, const T *, long, T *, long);
}  // namespace mg

using namespace mg;

namespace {

// NULL selects the HIP null stream, like any HIP API taking a stream.
hipStream_t as_stream(void *s) { return (hipStream_t)s; }

template <typename F>
int guarded(F &&f) {
  try {
    set_last_error("");
    return f();
  } catch (const Error &e) {
    set_last_error(e.msg);
    return e.code < 0 ? e.code - 100 : -100;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -199;
  }
}

int check_dims(size_t n) {
  MG_CHECK(n < (size_t)(1u << 31), "dimension too large");
  return 0;
}

template <typename T>
int gemm_ext(void *stream, int ta, int tb, size_t m, size_t n, size_t k, T alpha, const T *a,
             size_t lda, const T *b, size_t ldb, T beta, T *c, size_t ldc) {
  return guarded([&] {
    check_dims(m);
    check_dims(n);
    check_dims(k);
    View<const T> A = ta ? View<const T>{a, (long)lda, 1} : View<const T>{a, 1, (long)lda};
    View<const T> B = tb ? View<const T>{b, (long)ldb, 1} : View<const T>{b, 1, (long)ldb};
    gemm<T>(as_stream(stream), (long)m, (long)n, (long)k, alpha, A, B, beta,
            View<T>{c, 1, (long)ldc}, false);
    return 0;
  });
}

// Packed column-major LU through the row-major working copy.
template <typename T>
int getrf_ext(void *stream, size_t n_, T *a, size_t lda, int32_t *perm, int leaf, int *swaps,
              bool tournament = false) {
  return guarded([&] {
    check_dims(n_);
    const int n = (int)n_;
    hipStream_t s = as_stream(stream);
    if (n == 0) {
      if (swaps) *swaps = 0;
      return 0;
    }
    Workspace &ws = workspace(s);
    T *W = (T *)ws.get("ext.getrf.W", (size_t)n * n * sizeof(T));
    int *ipiv = (int *)ws.get("ext.getrf.ipiv", (size_t)n * 4);
    transpose<T>(s, n, n, a, (long)lda, W, n);
    LuOpts o;
    o.leaf = leaf > 0 ? leaf : 32;
    o.tournament = tournament;
    const T eps = sizeof(T) == 4 ? (T)1e-6f : (T)1e-9;
    getrf_rowmajor_perm<T>(s, n, W, n, ipiv, perm, eps, o);
    transpose<T>(s, n, n, W, n, a, (long)lda);
    std::vector<int> h(n);
    MG_HIP(hipMemcpyAsync(h.data(), ipiv, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    sync_and_check(s);
    std::vector<size_t> p(n);
    const int sw = ipiv_to_perm(n, h.data(), p.data());
    if (swaps) *swaps = sw;
    return 0;
  });
}

template <typename T>
int getrs_ext(void *stream, size_t n, size_t nrhs, const T *lu, size_t ldlu, const int32_t *perm,
              const T *b, size_t ldb, T *x, size_t ldx) {
  return guarded([&] {
    check_dims(n);
    check_dims(nrhs);
    hipStream_t s = as_stream(stream);
    // column-major packed factor: L(i,j) at lu[i + j*ld]
    gather_rows<T>(s, (int)n, (int)nrhs, b, (long)ldb, perm, x, (long)ldx);
    View<const T> LU{lu, 1, (long)ldlu};
    trsm_left<T>(s, true, true, (int)n, (int)nrhs, LU, View<T>{x, 1, (long)ldx});
    trsm_left<T>(s, false, false, (int)n, (int)nrhs, LU, View<T>{x, 1, (long)ldx});
    return 0;
  });
}

template <typename T>
int potrf_ext(void *stream, size_t n, T *a, size_t lda) {
  return guarded([&] {
    check_dims(n);
    hipStream_t s = as_stream(stream);
    int *info = (int *)workspace(s).get("ext.potrf.info", 4);
    potrf_lower<T>(s, (int)n, a, (long)lda, info);
    zero_upper<T>(s, (int)n, a, (long)lda);
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
    sync_and_check(s);
    return h != 0 ? -1 : 0;
  });
}

template <typename T>
int geqrf_ext(void *stream, size_t m, size_t n, T *a, size_t lda, T *q) {
  return guarded([&] {
    check_dims(m);
    check_dims(n);
    MG_CHECK(m >= n, "geqrf: needs m >= n");
    MG_CHECK(lda >= m, "geqrf: lda below m");
    hipStream_t s = as_stream(stream);
    geqrf<T>(s, (int)m, (int)n, a, (long)lda, q, sizeof(T) == 4 ? (T)1e-6f : (T)1e-9);
    sync_and_check(s);
    return 0;
  });
}

template <typename T>
int potrs_ext(void *stream, size_t n, size_t nrhs, const T *l, size_t ldl, const T *b, size_t ldb,
              T *x, size_t ldx) {
  return guarded([&] {
    check_dims(n);
    check_dims(nrhs);
    hipStream_t s = as_stream(stream);
    convert2d<T, T>(s, (int)n, (int)nrhs, b, (long)ldb, x, (long)ldx);
    potrs_lower<T>(s, (int)n, (int)nrhs, l, (long)ldl, x, (long)ldx);
    return 0;
  });
}

template <typename T>
int gesv_ext(void *stream, size_t n_, size_t nrhs, const T *a, size_t lda, const T *b, size_t ldb,
             T *x, size_t ldx) {
  return guarded([&] {
    check_dims(n_);
    check_dims(nrhs);
    const int n = (int)n_;
    hipStream_t s = as_stream(stream);
    Workspace &ws = workspace(s);
    T *W = (T *)ws.get("ext.gesv.W", (size_t)n * n * sizeof(T));
    int *ipiv = (int *)ws.get("ext.gesv.ipiv", (size_t)n * 4);
    int *perm = (int *)ws.get("ext.gesv.perm", (size_t)n * 4);
    transpose<T>(s, n, n, a, (long)lda, W, n);
    LuOpts o;
    const T eps = sizeof(T) == 4 ? (T)1e-6f : (T)1e-9;
    getrf_rowmajor_perm<T>(s, n, W, n, ipiv, perm, eps, o);
    getrs_rowmajor<T>(s, n, (int)nrhs, W, n, perm, b, (long)ldb, x, (long)ldx);
    return 0;
  });
}

template <typename T>
int posv_ext(void *stream, size_t n_, size_t nrhs, const T *a, size_t lda, const T *b, size_t ldb,
             T *x, size_t ldx) {
  return guarded([&] {
    check_dims(n_);
    check_dims(nrhs);
    const int n = (int)n_;
    hipStream_t s = as_stream(stream);
    Workspace &ws = workspace(s);
    T *L = (T *)ws.get("ext.posv.L", (size_t)n * n * sizeof(T));
    int *info = (int *)ws.get("ext.posv.info", 4);
    copy_lower<T>(s, n, a, (long)lda, L, n);
    potrf_lower<T>(s, n, L, n, info);
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
    sync_and_check(s);
    if (h != 0) return -1;
    convert2d<T, T>(s, n, (int)nrhs, b, (long)ldb, x, (long)ldx);
    potrs_lower<T>(s, n, (int)nrhs, L, n, x, (long)ldx);
    return 0;
  });
}

}  // namespace

extern "C" {

const char *mat_gpu_last_error(void) { return last_error(); }

int mat_gpu_timing_enable(int on) {
  return guarded([&] {
    timing_enable(on != 0);
    return 0;
  });
}

int mat_gpu_timing_query(const char *family, int64_t *launches, double *total_ms, double *work) {
  return guarded([&] {
    MG_CHECK(family != nullptr, "timing_query: family is NULL");
    long c = 0;
    timing_query(family, &c, total_ms, work);
    if (launches) *launches = c;
    return 0;
  });
}

int mat_gpu_stream_check(void *s) {
  return guarded([&] {
    sync_and_check(as_stream(s));
    return 0;
  });
}

int mat_gpu_gemm_f32(void *s, int ta, int tb, size_t m, size_t n, size_t k, float alpha,
                     const float *a, size_t lda, const float *b, size_t ldb, float beta, float *c,
                     size_t ldc) {
  return gemm_ext<float>(s, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mat_gpu_gemm_f64(void *s, int ta, int tb, size_t m, size_t n, size_t k, double alpha,
                     const double *a, size_t lda, const double *b, size_t ldb, double beta,
                     double *c, size_t ldc) {
  return gemm_ext<double>(s, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mat_gpu_getrf_f32(void *s, size_t n, float *a, size_t lda, int32_t *perm, int leaf, int *swaps) {
  return getrf_ext<float>(s, n, a, lda, perm, leaf, swaps);
}
int mat_gpu_getrf_f64(void *s, size_t n, double *a, size_t lda, int32_t *perm, int leaf,
                      int *swaps) {
  return getrf_ext<double>(s, n, a, lda, perm, leaf, swaps);
}
int mat_gpu_getrf_tournament_f32(void *s, size_t n, float *a, size_t lda, int32_t *perm,
                                 int *swaps) {
  return getrf_ext<float>(s, n, a, lda, perm, 32, swaps, true);
}
int mat_gpu_getrs_f32(void *s, size_t n, size_t nrhs, const float *lu, size_t ldlu,
                      const int32_t *perm, const float *b, size_t ldb, float *x, size_t ldx) {
  return getrs_ext<float>(s, n, nrhs, lu, ldlu, perm, b, ldb, x, ldx);
}
int mat_gpu_getrs_f64(void *s, size_t n, size_t nrhs, const double *lu, size_t ldlu,
                      const int32_t *perm, const double *b, size_t ldb, double *x, size_t ldx) {
  return getrs_ext<double>(s, n, nrhs, lu, ldlu, perm, b, ldb, x, ldx);
}
int mat_gpu_potrf_f32(void *s, size_t n, float *a, size_t lda) { return potrf_ext<float>(s, n, a, lda); }
int mat_gpu_potrf_f64(void *s, size_t n, double *a, size_t lda) {
  return potrf_ext<double>(s, n, a, lda);
}
int mat_gpu_geqrf_f32(void *s, size_t m, size_t n, float *a, size_t lda, float *q) {
  return geqrf_ext<float>(s, m, n, a, lda, q);
}
int mat_gpu_geqrf_f64(void *s, size_t m, size_t n, double *a, size_t lda, double *q) {
  return geqrf_ext<double>(s, m, n, a, lda, q);
}
int mat_gpu_potrs_f32(void *s, size_t n, size_t nrhs, const float *l, size_t ldl, const float *b,
                      size_t ldb, float *x, size_t ldx) {
  return potrs_ext<float>(s, n, nrhs, l, ldl, b, ldb, x, ldx);
}
int mat_gpu_potrs_f64(void *s, size_t n, size_t nrhs, const double *l, size_t ldl, const double *b,
                      size_t ldb, double *x, size_t ldx) {
  return potrs_ext<double>(s, n, nrhs, l, ldl, b, ldb, x, ldx);
}
int mat_gpu_gesv_f32(void *s, size_t n, size_t nrhs, const float *a, size_t lda, const float *b,
                     size_t ldb, float *x, size_t ldx) {
  return gesv_ext<float>(s, n, nrhs, a, lda, b, ldb, x, ldx);
}
int mat_gpu_gesv_f64(void *s, size_t n, size_t nrhs, const double *a, size_t lda, const double *b,
                     size_t ldb, double *x, size_t ldx) {
  return gesv_ext<double>(s, n, nrhs, a, lda, b, ldb, x, ldx);
}
int mat_gpu_posv_f32(void *s, size_t n, size_t nrhs, const float *a, size_t lda, const float *b,
                     size_t ldb, float *x, size_t ldx) {
  return posv_ext<float>(s, n, nrhs, a, lda, b, ldb, x, ldx);
}
int mat_gpu_posv_f64(void *s, size_t n, size_t nrhs, const double *a, size_t lda, const double *b,
                     size_t ldb, double *x, size_t ldx) {
  return posv_ext<double>(s, n, nrhs, a, lda, b, ldb, x, ldx);
}

int mat_gpu_gemm_lowp(void *s, int passes, size_t m, size_t n, size_t k, float alpha,
                      const float *a, size_t lda, const float *b, size_t ldb, float *c, size_t ldc) {
  return guarded([&] {
    check_dims(m);
    check_dims(n);
    check_dims(k);
    View<const float> A{a, 1, (long)lda}, B{b, 1, (long)ldb};
    View<float> Cv{c, 1, (long)ldc};
    if (passes == 8)
      gemm_fp8(as_stream(s), (long)m, (long)n, (long)k, alpha, A, B, Cv, false);
    else
      gemm_bf16x(as_stream(s), (long)m, (long)n, (long)k, alpha, A, B, Cv, passes);
    return 0;
  });
}

void mat_gpu_ir_defaults(mat_gpu_ir_opts *o, int spd) {
  if (!o) return;
  o->passes = spd ? 8 : 6;
  o->max_iter = 30;
  o->tol = 1e-14;
  o->leaf = 32;
  o->pivoting = spd ? 0 : 2;
}

int mat_gpu_gesv_mixed(void *s, size_t n, size_t nrhs, const double *a, size_t lda, const double *b,
                       size_t ldb, double *x, size_t ldx, const mat_gpu_ir_opts *o,
                       mat_gpu_ir_status *st) {
  return guarded([&] {
    check_dims(n);
    check_dims(nrhs);
    return gesv_mixed(as_stream(s), (int)n, (int)nrhs, a, (long)lda, b, (long)ldb, x, (long)ldx, o,
                      st);
  });
}

int mat_gpu_posv_mixed(void *s, size_t n, size_t nrhs, const double *a, size_t lda, const double *b,
                       size_t ldb, double *x, size_t ldx, const mat_gpu_ir_opts *o,
                       mat_gpu_ir_status *st) {
  return guarded([&] {
    check_dims(n);
    check_dims(nrhs);
    return posv_mixed(as_stream(s), (int)n, (int)nrhs, a, (long)lda, b, (long)ldb, x, (long)ldx, o,
                      st);
  });
}

}  // extern "C"
