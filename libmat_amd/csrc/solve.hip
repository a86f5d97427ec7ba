// Solves on top of the factorizations, and the mixed-precision solvers.
//
// getrs: mat_solve's second half (mat.h:4825-4830) for many right-hand
// sides at once: pb = b[p], unit-lower then upper substitution, here as
// blocked triangular solves with MFMA GEMM updates.
// potrs: mat_solve_spd's second half (mat.h:4862-4865): L y = b, L^T x = y.
//
// Mixed precision (extension, no counterpart in the reference):
//   gesv_mixed: fp32 working LU whose trailing updates run on BF16 MFMA
//   (split into 1, 3 or 6 bf16 products, fp32 accumulate), then iterative
//   refinement in which the residual r = b - A x (A in fp64) and the update
//   x += d are the only fp64 operations; d solves the fp32 factors.
//   posv_mixed: same with an fp32 Cholesky whose trailing updates run on
//   FP8 e4m3 MFMA with per-128x128-tile scales.
//   Both stop at backward error max_c ||r_c|| / (||A|| ||x_c|| + ||b_c||)
//   <= tol (infinity norms) and fall back to the plain fp64 factorization
//   when refinement stalls (two steps without halving the error, growth,
//   NaN, or the iteration cap).
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mat_gpu.h"
#include "ops.hpp"

namespace mg {

template <typename T>
int tri_block_size();
template <typename T>
void tri_invert_blocks(hipStream_t s, bool lower, bool unit, int n, View<const T> Tm, T *D);
template <typename T>
size_t tri_fold_size(bool lower, int n);
template <typename T>
void tri_fold_tiles(hipStream_t s, bool lower, int n, View<const T> Tm, const T *D, T *Tp);
template <typename T>
void tri_solve_blocked(hipStream_t s, bool lower, int n, int nrhs, View<const T> Tm, const T *D,
                       T *X, long ldx, const T *Tp);

// A pair of triangular factors with their inverted diagonal blocks, ready
// for repeated solves (one preparation per factorization).
template <typename T>
struct TriPair {
  int n = 0;
  View<const T> lo{}, up{};
  bool lo_unit = false, up_unit = false;
  T *dlo = nullptr, *dup = nullptr;
  T *flo = nullptr, *fup = nullptr;  // folded tiles (repeated solves only)

  // fold: also precompute D_r T(r,k) for every off-diagonal tile (one GEMM
  // per block row and triangle), which takes the diagonal-block product off
  // the critical path of every later solve; pays off from a few solves on
  void prepare(hipStream_t s, const char *tag, bool fold = false) {
    const int tb = tri_block_size<T>();
    const size_t blocks = (size_t)((n + tb - 1) / tb) * tb * tb;
    Workspace &ws = workspace(s);
    dlo = (T *)ws.get(std::string(tag) + ".dlo", blocks * sizeof(T));
    dup = (T *)ws.get(std::string(tag) + ".dup", blocks * sizeof(T));
    tri_invert_blocks<T>(s, true, lo_unit, n, lo, dlo);
    tri_invert_blocks<T>(s, false, up_unit, n, up, dup);
    flo = fup = nullptr;
    if (fold) {
      flo = (T *)ws.get(std::string(tag) + ".flo", tri_fold_size<T>(true, n) * sizeof(T));
      fup = (T *)ws.get(std::string(tag) + ".fup", tri_fold_size<T>(false, n) * sizeof(T));
      tri_fold_tiles<T>(s, true, n, lo, dlo, flo);
      tri_fold_tiles<T>(s, false, n, up, dup, fup);
    }
  }
  // X := (lo * up)^-1 X
  void solve(hipStream_t s, int nrhs, T *X, long ldx) const {
    if (nrhs <= 64) {
      tri_solve_blocked<T>(s, true, n, nrhs, lo, dlo, X, ldx, flo);
      tri_solve_blocked<T>(s, false, n, nrhs, up, dup, X, ldx, fup);
    } else {
      trsm_left<T>(s, true, lo_unit, n, nrhs, lo, View<T>{X, 1, ldx});
      trsm_left<T>(s, false, up_unit, n, nrhs, up, View<T>{X, 1, ldx});
    }
  }
};

void rbt_matrix(hipStream_t, int, int, const double *, long, const double *, const double *, float *,
                long);
void rbt_vec_t(hipStream_t, int, int, int, const double *, const double *, long, float *, long);
void rbt_vec_n(hipStream_t, int, int, int, const double *, const float *, long, double *, long, bool);
void getrf_nopiv_rowmajor(hipStream_t, int, float *, long, float, Mul);

// Butterfly diagonals exp(rho/10), rho ~ U[-1/2, 1/2], from a fixed seed
// (splitmix64), laid out [r0 (h), r1 (h), s0, s1, t0, t1 (q each)].
static std::vector<double> butterfly_diagonals(int n4, uint64_t seed) {
  std::vector<double> v((size_t)2 * n4);
  uint64_t x = seed;
  for (auto &e : v) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * 0x1.0p-53;  // [0, 1)
    e = std::exp((u - 0.5) / 10.0);
  }
  return v;
}

template <typename T>
TriPair<T> lu_pair(int n, const T *W, long ld) {
  TriPair<T> p;
  p.n = n;
  p.lo = View<const T>{W, ld, 1};
  p.up = p.lo;
  p.lo_unit = true;
  return p;
}

template <typename T>
TriPair<T> chol_pair(int n, const T *L, long ld) {
  TriPair<T> p;
  p.n = n;
  p.lo = View<const T>{L, 1, ld};
  p.up = p.lo.t();
  p.lo_unit = false;
  return p;
}

template <typename T>
void getrs_rowmajor(hipStream_t s, int n, int nrhs, const T *W, long ld, const int *perm,
                    const T *B, long ldb, T *X, long ldx) {
  gather_rows<T>(s, n, nrhs, B, ldb, perm, X, ldx);
  TriPair<T> tp = lu_pair<T>(n, W, ld);
  if (nrhs <= 64) tp.prepare(s, "getrs");
  tp.solve(s, nrhs, X, ldx);
}

template <typename T>
void potrs_lower(hipStream_t s, int n, int nrhs, const T *L, long ld, T *X, long ldx) {
  TriPair<T> tp = chol_pair<T>(n, L, ld);
  if (nrhs <= 64) tp.prepare(s, "potrs");
  tp.solve(s, nrhs, X, ldx);
}

template void getrs_rowmajor<float>(hipStream_t, int, int, const float *, long, const int *,
                                    const float *, long, float *, long);
template void getrs_rowmajor<double>(hipStream_t, int, int, const double *, long, const int *,
                                     const double *, long, double *, long);
template void potrs_lower<float>(hipStream_t, int, int, const float *, long, float *, long);
template void potrs_lower<double>(hipStream_t, int, int, const double *, long, double *, long);

// ------------------------------------------------------ fp64 residual ----
// R = B - A X, A n x n column-major fp64, up to 16 right-hand sides per
// pass.  Split-K: workgroup (rb, kc) owns 512 rows and one 256-deep K
// chunk; the chunk of X is staged in LDS (broadcast reads), each thread
// streams two rows of A (i and i+256: each load coalesced across the
// wavefront) into 2 x 16 partial sums, two k steps at a time (their A
// values and X rows loaded before the FMAs).  Two rows per thread halve
// the LDS reads per element of A (measured at n = 16384, 16 right-hand
// sides: 402 us against 481 us with one row per thread; four rows per
// thread ran slower, fewer waves hiding the latencies).  Partials are
// reduced in a fixed order (deterministic); each sum runs over k in order.
constexpr int RES_NR = 16;
constexpr int RES_KC = 256;
constexpr int RES_RPT = 2;  // rows per thread
constexpr int RES_U = 2;    // k steps per round

__global__ __launch_bounds__(256) void residual_partial_kernel(int n, int nrhs, const double *A,
                                                               long lda, const double *X, long ldx,
                                                               double *part) {
  __shared__ double Xs[RES_KC * RES_NR];  // [k][c]
  const int i0 = blockIdx.x * (256 * RES_RPT) + threadIdx.x;
  const int k0 = blockIdx.y * RES_KC;
  const int kn = min(RES_KC, n - k0);
  for (int idx = threadIdx.x; idx < RES_KC * RES_NR; idx += 256) {
    const int kk = idx % RES_KC, c = idx / RES_KC;
    Xs[kk * RES_NR + c] = (kk < kn && c < nrhs) ? X[k0 + kk + (long)c * ldx] : 0.0;
  }
  __syncthreads();
  if (i0 >= n) return;
  double acc[RES_RPT][RES_NR];
#pragma unroll
  for (int r = 0; r < RES_RPT; ++r)
#pragma unroll
    for (int c = 0; c < RES_NR; ++c) acc[r][c] = 0.0;
  // rows past n read row i0 (in range) and are not stored
  long roff[RES_RPT];
#pragma unroll
  for (int r = 0; r < RES_RPT; ++r) roff[r] = i0 + 256 * r < n ? 256 * r : 0;
  const double *a = A + i0 + (long)k0 * lda;
  int kk = 0;
  for (; kk + RES_U <= kn; kk += RES_U) {
    double v[RES_U][RES_RPT];
#pragma unroll
    for (int u = 0; u < RES_U; ++u)
#pragma unroll
      for (int r = 0; r < RES_RPT; ++r) v[u][r] = a[(long)(kk + u) * lda + roff[r]];
    double x[RES_U][RES_NR];
#pragma unroll
    for (int u = 0; u < RES_U; ++u)
#pragma unroll
      for (int c2 = 0; c2 < RES_NR / 2; ++c2) {
        const double2 q = *reinterpret_cast<const double2 *>(Xs + (kk + u) * RES_NR + 2 * c2);
        x[u][2 * c2] = q.x;
        x[u][2 * c2 + 1] = q.y;
      }
#pragma unroll
    for (int u = 0; u < RES_U; ++u)
#pragma unroll
      for (int r = 0; r < RES_RPT; ++r)
#pragma unroll
        for (int c = 0; c < RES_NR; ++c) acc[r][c] = fma(v[u][r], x[u][c], acc[r][c]);
  }
  for (; kk < kn; ++kk) {
#pragma unroll
    for (int r = 0; r < RES_RPT; ++r) {
      const double v = a[(long)kk * lda + roff[r]];
#pragma unroll
      for (int c = 0; c < RES_NR; ++c) acc[r][c] = fma(v, Xs[kk * RES_NR + c], acc[r][c]);
    }
  }
#pragma unroll
  for (int r = 0; r < RES_RPT; ++r) {
    if (i0 + 256 * r >= n) continue;
    double *pp = part + ((long)blockIdx.y * RES_NR) * n + i0 + 256 * r;
#pragma unroll
    for (int c = 0; c < RES_NR; ++c)
      if (c < nrhs) pp[(long)c * n] = acc[r][c];
  }
}

__global__ void residual_reduce_kernel(int n, int nrhs, int nchunks, const double *part,
                                       const double *B, long ldb, double *R, long ldr) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * nrhs) return;
  const int i = (int)(idx % n), c = (int)(idx / n);
  double s = 0.0;
  for (int q = 0; q < nchunks; ++q) s += part[((long)q * RES_NR + c) * n + i];
  R[i + (long)c * ldr] = B[i + (long)c * ldb] - s;
}

void residual_f64(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *X,
                  long ldx, const double *B, long ldb, double *R, long ldr) {
  const int nchunks = ceil_div(n, RES_KC);
  double *part = (double *)workspace(s).get("res.part", (size_t)nchunks * RES_NR * n * 8);
  for (int c0 = 0; c0 < nrhs; c0 += RES_NR) {
    const int nr = std::min(RES_NR, nrhs - c0);
    hipLaunchKernelGGL(residual_partial_kernel, dim3(ceil_div(n, 256 * RES_RPT), nchunks),
                       dim3(256), 0, s, n, nr, A, lda, X + (long)c0 * ldx, ldx, part);
    hipLaunchKernelGGL(residual_reduce_kernel, dim3(ceil_div((size_t)n * nr, 256)), dim3(256), 0, s,
                       n, nr, nchunks, part, B + (long)c0 * ldb, ldb, R + (long)c0 * ldr, ldr);
  }
  MG_HIP(hipGetLastError());
}

// ------------------------------------------------------------- norms -----
// out[c] = max_i |X(i,c)| (non-negative doubles order like their bits).
// max that keeps NaN (fmax drops it, which would hide a NaN solution); as a
// 64-bit pattern a NaN also orders above +inf for the atomicMax below
__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

__global__ void col_amax_kernel(int n, const double *X, long ldx, unsigned long long *out) {
  const int c = blockIdx.y;
  double m = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    m = nan_max(m, fabs(X[i + (long)c * ldx]));
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out + c, (unsigned long long)__double_as_longlong(red[0]));
}

// partial row sums of |A| over 256-column chunks: part[chunk][i]
__global__ __launch_bounds__(256) void row_abs_partial_kernel(int n, const double *A, long lda,
                                                              double *part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * 256, j1 = min(n, j0 + 256);
  if (i >= n) return;
  double s = 0.0;
  for (int j = j0; j < j1; ++j) s += fabs(A[i + (long)j * lda]);
  part[(long)blockIdx.y * n + i] = s;
}

__global__ void row_abs_max_kernel(int n, int nchunks, const double *part,
                                   unsigned long long *out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < n)
    for (int q = 0; q < nchunks; ++q) s += part[(long)q * n + i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

static void col_amax(hipStream_t s, int n, int nrhs, const double *X, long ldx,
                     unsigned long long *out) {
  hipLaunchKernelGGL(col_amax_kernel, dim3(std::min(ceil_div(n, 256), 64u), nrhs), dim3(256), 0, s,
                     n, X, ldx, out);
  MG_HIP(hipGetLastError());
}

// ||A||_inf (max row sum), fp64, synchronous.
double norm_inf_f64(hipStream_t s, int n, const double *A, long lda) {
  const int nchunks = ceil_div(n, 256);
  Workspace &ws = workspace(s);
  unsigned long long *d = (unsigned long long *)ws.get("norm.A", 8);
  double *part = (double *)ws.get("norm.part", (size_t)nchunks * n * 8);
  fill_u32(s, (unsigned *)d, 2, 0u);
  hipLaunchKernelGGL(row_abs_partial_kernel, dim3(ceil_div(n, 256), nchunks), dim3(256), 0, s, n, A,
                     lda, part);
  hipLaunchKernelGGL(row_abs_max_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, nchunks, part, d);
  MG_HIP(hipGetLastError());
  unsigned long long h = 0;
  MG_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, s));
  sync_and_check(s);
  double v;
  memcpy(&v, &h, 8);
  return v;
}

// x += (double) d   (column-major blocks)
__global__ void axpy_f32_to_f64_kernel(int n, int nrhs, const float *D, long ldd, double *X,
                                       long ldx) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * nrhs) return;
  const int i = (int)(idx % n), c = (int)(idx / n);
  X[i + (long)c * ldx] += (double)D[i + (long)c * ldd];
}

template <typename Ti, typename To>
__global__ void transpose_convert_kernel(int rows, int cols, const Ti *src, long lds, To *dst,
                                         long ldd) {
  __shared__ Ti tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int i = bx + tx, j = by + yy;
    if (i < rows && j < cols) tile[yy][tx] = src[i + (long)j * lds];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int j = by + tx, i = bx + yy;
    if (i < rows && j < cols) dst[j + (long)i * ldd] = (To)tile[tx][yy];
  }
}

template <typename Ti, typename To>
static void transpose_convert(hipStream_t s, int rows, int cols, const Ti *src, long lds, To *dst,
                              long ldd) {
  hipLaunchKernelGGL((transpose_convert_kernel<Ti, To>), dim3(ceil_div(rows, 32), ceil_div(cols, 32)),
                     dim3(256), 0, s, rows, cols, src, lds, dst, ldd);
  MG_HIP(hipGetLastError());
}

// The backward error's column maxima go to the host through a mailbox in
// pinned memory: one workgroup copies them, and the multi-workgroup
// kernels' error word, with system-scope stores, then raises a tag; the
// host polls the tag.  (A copy to pageable memory and a stream
// synchronization cost ~70 us per refinement step: a staged copy and the
// wake-up of the waiting thread.)
__global__ void berr_publish_kernel(const unsigned long long *dev, int count, const int *err,
                                    unsigned long long *mbox, unsigned long long tag) {
  for (int i = threadIdx.x; i < count; i += blockDim.x)
    __hip_atomic_store(mbox + i, dev[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  if (threadIdx.x == 0)
    __hip_atomic_store(mbox + count, (unsigned long long)(unsigned)*err, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __hip_atomic_store(mbox + count + 1, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

static std::atomic<unsigned long long> g_mbox_tag{0};

// Backward error per column, max over columns, from device vectors.
struct BerrEval {
  hipStream_t s;
  int n, nrhs;
  double normA;
  unsigned long long *dev;   // [3][nrhs]: |r|, |x|, |b|
  unsigned long long *mbox;  // pinned: [3][nrhs] copies, the error word, the tag
  int *err;

  BerrEval(hipStream_t s_, int n_, int nrhs_, double normA_, const double *B, long ldb)
      : s(s_), n(n_), nrhs(nrhs_), normA(normA_) {
    Workspace &ws = workspace(s);
    dev = (unsigned long long *)ws.get("berr", (size_t)3 * nrhs * 8);
    mbox = (unsigned long long *)ws.get_host("berr.mbox", ((size_t)3 * nrhs + 2) * 8);
    err = sync_error_word(s);
    fill_u32(s, (unsigned *)dev, 6L * nrhs, 0u);
    col_amax(s, n, nrhs, B, ldb, dev + 2 * nrhs);
  }
  double eval(const double *R, long ldr, const double *X, long ldx) {
    fill_u32(s, (unsigned *)dev, 4L * nrhs, 0u);
    col_amax(s, n, nrhs, R, ldr, dev);
    col_amax(s, n, nrhs, X, ldx, dev + nrhs);
    const int count = 3 * nrhs;
    const unsigned long long tag = ++g_mbox_tag;
    hipLaunchKernelGGL(berr_publish_kernel, dim3(1), dim3(64), 0, s, dev, count, err, mbox, tag);
    MG_HIP(hipGetLastError());
    // poll (a refinement step takes a few ms); past 50 ms, block on the
    // stream instead, which also reports any error of the queued work
    unsigned long long *tagp = mbox + count + 1;
    const auto t0 = std::chrono::steady_clock::now();
    for (long spin = 0; __atomic_load_n(tagp, __ATOMIC_ACQUIRE) != tag; ++spin) {
      if ((spin & 1023) == 1023 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
        MG_HIP(hipStreamSynchronize(s));
        MG_CHECK(__atomic_load_n(tagp, __ATOMIC_ACQUIRE) == tag,
                 "backward error: the mailbox was not written");
        break;
      }
    }
    if (__atomic_load_n(mbox + count, __ATOMIC_ACQUIRE) != 0) {
      MG_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
      MG_HIP(hipStreamSynchronize(s));
      throw Error{-5, "a multi-workgroup kernel timed out waiting for another workgroup"};
    }
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c) {
      const unsigned long long hr = __atomic_load_n(mbox + c, __ATOMIC_RELAXED);
      const unsigned long long hx = __atomic_load_n(mbox + nrhs + c, __ATOMIC_RELAXED);
      const unsigned long long hb = __atomic_load_n(mbox + 2 * nrhs + c, __ATOMIC_RELAXED);
      double r, x, b;
      memcpy(&r, &hr, 8);
      memcpy(&x, &hx, 8);
      memcpy(&b, &hb, 8);
      const double den = normA * x + b;
      double e;
      if (!(den == den) || !(r == r)) e = NAN;  // NaN anywhere: never converged
      else if (den > 0) e = r / den;
      else e = r > 0 ? INFINITY : 0.0;
      if (!(e <= worst)) worst = e;  // NaN propagates
    }
    return worst;
  }
};

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Plain fp64 partial-pivot LU solve (device pointers, column-major).
int gesv_f64(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *B, long ldb,
             double *X, long ldx, int leaf) {
  Workspace &ws = workspace(s);
  double *W = (double *)ws.get("gesv64.W", (size_t)n * n * 8);
  int *ipiv = (int *)ws.get("gesv64.ipiv", (size_t)n * 4);
  int *perm = (int *)ws.get("gesv64.perm", (size_t)n * 4);
  transpose<double>(s, n, n, A, lda, W, n);
  LuOpts o;
  o.leaf = leaf;
  getrf_rowmajor_perm<double>(s, n, W, n, ipiv, perm, 1e-9, o);
  getrs_rowmajor<double>(s, n, nrhs, W, n, perm, B, ldb, X, ldx);
  return 0;
}

int gesv_mixed(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *B, long ldb,
               double *X, long ldx, const mat_gpu_ir_opts *opt, mat_gpu_ir_status *st) {
  const auto t0 = std::chrono::steady_clock::now();
  mat_gpu_ir_opts o{};
  o.passes = 6;
  o.max_iter = 30;
  o.tol = 1e-14;
  o.leaf = 32;
  o.pivoting = 2;
  if (opt) o = *opt;
  MG_CHECK(o.passes == 0 || o.passes == 1 || o.passes == 3 || o.passes == 6,
           "gesv_mixed: passes must be 0 (fp32 MFMA), 1, 3 or 6");
  MG_CHECK(o.pivoting >= 0 && o.pivoting <= 2,
           "gesv_mixed: pivoting must be 0 (partial), 1 (rbt) or 2 (tournament)");
  const bool rbt = o.pivoting == 1;
  const int n4 = rbt ? (n + 3) / 4 * 4 : n;
  Workspace &ws = workspace(s);
  float *W = (float *)ws.get("gesvm.W", (size_t)n4 * n4 * 4);
  int *ipiv = (int *)ws.get("gesvm.ipiv", (size_t)n4 * 4);
  int *perm = (int *)ws.get("gesvm.perm", (size_t)n4 * 4);
  double *R = (double *)ws.get("gesvm.R", (size_t)n * nrhs * 8);
  float *D2 = (float *)ws.get("gesvm.D2", (size_t)n4 * nrhs * 4);
  double *ubf = nullptr, *vbf = nullptr;
  const Mul mul = o.passes == 0 ? Mul::Native : (Mul)o.passes;

  const double normA = norm_inf_f64(s, n, A, lda);
  TriPair<float> tp;
  if (rbt) {
    ubf = (double *)ws.get("gesvm.ubf", (size_t)2 * n4 * 8);
    vbf = (double *)ws.get("gesvm.vbf", (size_t)2 * n4 * 8);
    const std::vector<double> hu = butterfly_diagonals(n4, 0x1234567ull + (uint64_t)n4);
    const std::vector<double> hv = butterfly_diagonals(n4, 0x7654321ull + (uint64_t)n4);
    MG_HIP(hipMemcpyAsync(ubf, hu.data(), hu.size() * 8, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(vbf, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, s));
    // W (column-major) = U^T A V, i.e. row-major W holds M = A_r^T;  M = L U
    // gives A_r = U^T L^T: lower non-unit then upper unit, both read through
    // the column-major view of W.
    rbt_matrix(s, n, n4, A, lda, ubf, vbf, W, n4);
    getrf_nopiv_rowmajor(s, n4, W, n4, 1e-6f, mul);
    tp.n = n4;
    tp.lo = View<const float>{W, 1, n4};
    tp.up = tp.lo;
    tp.lo_unit = false;
    tp.up_unit = true;
    sync_and_check(s);  // host vectors above must outlive the copies
  } else {
    transpose_convert<double, float>(s, n, n, A, lda, W, n);
    LuOpts lo;
    lo.leaf = o.leaf;
    lo.mul = mul;
    lo.tournament = o.pivoting == 2;
    getrf_rowmajor_perm<float>(s, n, W, n, ipiv, perm, 1e-6f, lo);
    tp = lu_pair<float>(n, W, n);
  }
  // folded tiles need bounded diagonal-block inverses: partial pivoting
  // gives them, the pivot-free butterfly path may not
  tp.prepare(s, "gesvm", !rbt);
  sync_and_check(s);
  const double t_factor = ms_since(t0);

  // d = A^-1 r through the fp32 factors (r = B on the first pass)
  auto correct = [&](const double *Rm, long ldr, bool assign) {
    if (rbt) {
      rbt_vec_t(s, n, n4, nrhs, ubf, Rm, ldr, D2, n4);
      tp.solve(s, nrhs, D2, n4);
      rbt_vec_n(s, n, n4, nrhs, vbf, D2, n4, X, ldx, assign);
    } else {
      gather_rows_cvt<double, float>(s, n, nrhs, Rm, ldr, perm, D2, n);
      tp.solve(s, nrhs, D2, n);
      if (assign) convert2d<float, double>(s, n, nrhs, D2, n, X, ldx);
      else
        hipLaunchKernelGGL(axpy_f32_to_f64_kernel, dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256),
                           0, s, n, nrhs, D2, n, X, ldx);
    }
    MG_HIP(hipGetLastError());
  };

  BerrEval be(s, n, nrhs, normA, B, ldb);
  correct(B, ldb, true);
  int it = 0;
  double berr = INFINITY, prev = INFINITY, prev2 = INFINITY;
  bool ok = false;
  while (true) {
    residual_f64(s, n, nrhs, A, lda, X, ldx, B, ldb, R, n);
    berr = be.eval(R, n, X, ldx);
    if (berr <= o.tol) {
      ok = true;
      break;
    }
    const bool stalled = !(berr == berr) || it >= o.max_iter ||
                         (it >= 2 && berr > 0.5 * prev2) || (it >= 1 && berr > 2.0 * prev);
    if (stalled) break;
    prev2 = prev;
    prev = berr;
    correct(R, n, false);
    ++it;
  }
  if (st) {
    st->iterations = it;
    st->fell_back = 0;
    st->backward_error = berr;
    st->factor_ms = t_factor;
    st->precision = o.passes;
  }
  if (!ok) {
    // refinement stalled: plain fp64 factorization
    gesv_f64(s, n, nrhs, A, lda, B, ldb, X, ldx, o.leaf);
    residual_f64(s, n, nrhs, A, lda, X, ldx, B, ldb, R, n);
    berr = be.eval(R, n, X, ldx);
    if (st) {
      st->fell_back = 1;
      st->backward_error = berr;
    }
  }
  if (st) st->total_ms = ms_since(t0);
  sync_and_check(s);
  return 0;
}

// Plain fp64 Cholesky solve (device pointers); returns -1 if not SPD.
int posv_f64(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *B, long ldb,
             double *X, long ldx) {
  Workspace &ws = workspace(s);
  double *L = (double *)ws.get("posv64.L", (size_t)n * n * 8);
  int *info = (int *)ws.get("posv64.info", 4);
  copy_lower<double>(s, n, A, lda, L, n);
  potrf_lower<double>(s, n, L, n, info);
  int h = 0;
  MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
  sync_and_check(s);
  if (h != 0) return -1;
  convert2d<double, double>(s, n, nrhs, B, ldb, X, ldx);
  potrs_lower<double>(s, n, nrhs, L, n, X, ldx);
  return 0;
}

int posv_mixed(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *B, long ldb,
               double *X, long ldx, const mat_gpu_ir_opts *opt, mat_gpu_ir_status *st) {
  const auto t0 = std::chrono::steady_clock::now();
  mat_gpu_ir_opts o{};
  o.passes = 8;
  o.max_iter = 30;
  o.tol = 1e-14;
  o.leaf = 32;
  o.pivoting = 0;
  if (opt) o = *opt;
  MG_CHECK(o.passes == 0 || o.passes == 1 || o.passes == 3 || o.passes == 6 || o.passes == 8,
           "posv_mixed: passes must be 0 (fp32), 1/3/6 (bf16) or 8 (fp8)");
  Workspace &ws = workspace(s);
  float *L = (float *)ws.get("posvm.L", (size_t)n * n * 4);
  int *info = (int *)ws.get("posvm.info", 4);
  double *R = (double *)ws.get("posvm.R", (size_t)n * nrhs * 8);
  float *D = (float *)ws.get("posvm.D", (size_t)n * nrhs * 4);

  const double normA = norm_inf_f64(s, n, A, lda);
  BerrEval be(s, n, nrhs, normA, B, ldb);
  TriPair<float> tp = chol_pair<float>(n, L, n);
  // precision ladder: the requested updates, then (after fp8) 3-pass bf16,
  // then the plain fp64 solver below
  const int levels[2] = {o.passes, o.passes == 8 ? 3 : -1};
  int it = 0, used = o.passes;
  double berr = INFINITY, t_factor = 0.0;
  bool ok = false;
  for (int li = 0; li < 2 && !ok && levels[li] >= 0; ++li) {
    const int ps = levels[li];
    convert2d<double, float>(s, n, n, A, lda, L, n);
    potrf_lower<float>(s, n, L, n, info, ps == 0 ? Mul::Native : (Mul)ps);
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, info, 4, hipMemcpyDeviceToHost, s));
    sync_and_check(s);
    if (li == 0) t_factor = ms_since(t0);
    if (h != 0) continue;  // not positive definite at this precision
    used = ps;
    tp.prepare(s, "posvm", true);
    convert2d<double, float>(s, n, nrhs, B, ldb, D, n);
    tp.solve(s, nrhs, D, n);
    convert2d<float, double>(s, n, nrhs, D, n, X, ldx);
    double prev = INFINITY, prev2 = INFINITY;
    int lit = 0;
    while (true) {
      residual_f64(s, n, nrhs, A, lda, X, ldx, B, ldb, R, n);
      berr = be.eval(R, n, X, ldx);
      if (berr <= o.tol) {
        ok = true;
        break;
      }
      const bool stalled = !(berr == berr) || lit >= o.max_iter ||
                           (lit >= 2 && berr > 0.5 * prev2) || (lit >= 1 && berr > 2.0 * prev);
      if (stalled) break;
      prev2 = prev;
      prev = berr;
      convert2d<double, float>(s, n, nrhs, R, n, D, n);
      tp.solve(s, nrhs, D, n);
      hipLaunchKernelGGL(axpy_f32_to_f64_kernel, dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256), 0,
                         s, n, nrhs, D, n, X, ldx);
      MG_HIP(hipGetLastError());
      ++lit;
      ++it;
    }
  }
  if (st) {
    st->iterations = it;
    st->fell_back = 0;
    st->backward_error = berr;
    st->factor_ms = t_factor;
    st->precision = used;
  }
  int rc = 0;
  if (!ok) {
    rc = posv_f64(s, n, nrhs, A, lda, B, ldb, X, ldx);
    if (rc == 0) {
      residual_f64(s, n, nrhs, A, lda, X, ldx, B, ldb, R, n);
      berr = be.eval(R, n, X, ldx);
    }
    if (st) {
      st->fell_back = 1;
      st->backward_error = berr;
    }
  }
  if (st) st->total_ms = ms_since(t0);
  sync_and_check(s);
  return rc;
}

}  // namespace mg
