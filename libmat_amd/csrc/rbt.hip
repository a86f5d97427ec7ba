// Recursive random butterfly transform (RBT) and pivot-free LU, used by the
// mixed-precision general solver as an alternative to partial pivoting.
//
// A butterfly of order 2m with random diagonals a, b (entries exp(rho/10),
// rho ~ U[-1/2, 1/2]) is B = (1/sqrt 2) [[diag a, diag b], [diag a, -diag b]].
// Depth 2: U = diag(B1, B2) B0 (B0 of order n, B1/B2 of order n/2), likewise
// V.  The solver factors A_r = U^T A V without pivoting (Parker 1995;
// Baboulin, Dongarra, Herrmann, Tomov, "Accelerating linear system solutions
// using randomization techniques", ACM TOMS 2013), then
//   A x = r  <=>  A_r y = U^T r,  x = V y.
// Refinement residuals are always taken against the original A in fp64.
//
// The transform touches each element once: a thread owns the 4 x 4 group
// {i, i+q, i+h, i+h+q} x {j, j+q, j+h, j+h+q} (q = n/4, h = n/2), applies
// the two row stages (U^T) and the two column stages (V), and writes fp32.
// n is padded to a multiple of 4 with an identity block.
#include "ops.hpp"

namespace mg {

constexpr double RSQRT2 = 0.70710678118654752440;

// Butterfly diagonals for one side: [r0 (h), r1 (h), s0 (q), s1 (q), t0 (q), t1 (q)]
struct Bfly {
  const double *r0, *r1, *s0, *s1, *t0, *t1;
};

__device__ inline Bfly bfly_at(const double *p, int n4) {
  const int h = n4 / 2, q = n4 / 4;
  Bfly b;
  b.r0 = p;
  b.r1 = p + h;
  b.s0 = p + 2 * h;
  b.s1 = b.s0 + q;
  b.t0 = b.s1 + q;
  b.t1 = b.t0 + q;
  return b;
}

// y <- B0^T diag(B1^T, B2^T) y on the group (y0..y3) = indices (i, i+q, i+h, i+h+q)
__device__ inline void apply_t(const Bfly &b, int i, int q, double y[4]) {
  const double z0 = b.s0[i] * (y[0] + y[1]) * RSQRT2;
  const double z1 = b.s1[i] * (y[0] - y[1]) * RSQRT2;
  const double z2 = b.t0[i] * (y[2] + y[3]) * RSQRT2;
  const double z3 = b.t1[i] * (y[2] - y[3]) * RSQRT2;
  y[0] = b.r0[i] * (z0 + z2) * RSQRT2;
  y[2] = b.r1[i] * (z0 - z2) * RSQRT2;
  y[1] = b.r0[i + q] * (z1 + z3) * RSQRT2;
  y[3] = b.r1[i + q] * (z1 - z3) * RSQRT2;
}

// y <- diag(B1, B2) B0 y on the same group
__device__ inline void apply_n(const Bfly &b, int i, int q, double y[4]) {
  const double z0 = (b.r0[i] * y[0] + b.r1[i] * y[2]) * RSQRT2;
  const double z2 = (b.r0[i] * y[0] - b.r1[i] * y[2]) * RSQRT2;
  const double z1 = (b.r0[i + q] * y[1] + b.r1[i + q] * y[3]) * RSQRT2;
  const double z3 = (b.r0[i + q] * y[1] - b.r1[i + q] * y[3]) * RSQRT2;
  y[0] = (b.s0[i] * z0 + b.s1[i] * z1) * RSQRT2;
  y[1] = (b.s0[i] * z0 - b.s1[i] * z1) * RSQRT2;
  y[2] = (b.t0[i] * z2 + b.t1[i] * z3) * RSQRT2;
  y[3] = (b.t0[i] * z2 - b.t1[i] * z3) * RSQRT2;
}

// Out (fp32, column-major n4 x n4, ld) = U^T A_pad V; A fp64 column-major n x n.
__global__ __launch_bounds__(256) void rbt_matrix_kernel(int n, int n4, const double *A, long lda,
                                                         const double *ubf, const double *vbf,
                                                         float *Out, long ldo) {
  const int q = n4 / 4;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= q || j >= q) return;
  const Bfly U = bfly_at(ubf, n4), V = bfly_at(vbf, n4);
  double a[4][4];  // a[row r][col c], rows i + r*q, cols j + c*q
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int col = j + c * q;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i + r * q;
      double v;
      if (row < n && col < n) v = A[row + (long)col * lda];
      else v = (row == col) ? 1.0 : 0.0;
      a[r][c] = v;
    }
  }
  // rows: U^T on each column of the group
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double y[4] = {a[0][c], a[1][c], a[2][c], a[3][c]};
    apply_t(U, i, q, y);
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r][c] = y[r];
  }
  // columns: x V = (V^T x^T)^T on each row of the group
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double y[4] = {a[r][0], a[r][1], a[r][2], a[r][3]};
    apply_t(V, j, q, y);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[r][c] = y[c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) Out[(i + r * q) + (long)(j + c * q) * ldo] = (float)a[r][c];
}

// Y (fp32 n4 x nrhs) = U^T R_pad  (R fp64 n x nrhs, zero padded)
__global__ void rbt_vec_t_kernel(int n, int n4, int nrhs, const double *ubf, const double *R,
                                 long ldr, float *Y, long ldy) {
  const int q = n4 / 4;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= q || c >= nrhs) return;
  const Bfly U = bfly_at(ubf, n4);
  double y[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = i + r * q;
    y[r] = row < n ? R[row + (long)c * ldr] : 0.0;
  }
  apply_t(U, i, q, y);
#pragma unroll
  for (int r = 0; r < 4; ++r) Y[(i + r * q) + (long)c * ldy] = (float)y[r];
}

// X (fp64 n x nrhs) += V Y (Y fp32 n4 x nrhs), or X = V Y when assign
__global__ void rbt_vec_n_kernel(int n, int n4, int nrhs, const double *vbf, const float *Y,
                                 long ldy, double *X, long ldx, int assign) {
  const int q = n4 / 4;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= q || c >= nrhs) return;
  const Bfly V = bfly_at(vbf, n4);
  double y[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) y[r] = (double)Y[(i + r * q) + (long)c * ldy];
  apply_n(V, i, q, y);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = i + r * q;
    if (row < n) {
      double *x = X + row + (long)c * ldx;
      *x = assign ? y[r] : *x + y[r];
    }
  }
}

void rbt_matrix(hipStream_t s, int n, int n4, const double *A, long lda, const double *ubf,
                const double *vbf, float *Out, long ldo) {
  const int q = n4 / 4;
  hipLaunchKernelGGL(rbt_matrix_kernel, dim3(ceil_div(q, 64), ceil_div(q, 4)), dim3(256), 0, s, n,
                     n4, A, lda, ubf, vbf, Out, ldo);
  MG_HIP(hipGetLastError());
}

void rbt_vec_t(hipStream_t s, int n, int n4, int nrhs, const double *ubf, const double *R, long ldr,
               float *Y, long ldy) {
  hipLaunchKernelGGL(rbt_vec_t_kernel, dim3(ceil_div(n4 / 4, 256), nrhs), dim3(256), 0, s, n, n4,
                     nrhs, ubf, R, ldr, Y, ldy);
  MG_HIP(hipGetLastError());
}

void rbt_vec_n(hipStream_t s, int n, int n4, int nrhs, const double *vbf, const float *Y, long ldy,
               double *X, long ldx, bool assign) {
  hipLaunchKernelGGL(rbt_vec_n_kernel, dim3(ceil_div(n4 / 4, 256), nrhs), dim3(256), 0, s, n, n4,
                     nrhs, vbf, Y, ldy, X, ldx, assign ? 1 : 0);
  MG_HIP(hipGetLastError());
}

// ----------------------------------------------------- pivot-free LU ----
// Diagonal block (w <= 64) of the row-major working matrix, right-looking in
// LDS; a pivot below eps skips its step (mat.h:4234 rule).
constexpr int NP_LEAF = 64;

__global__ __launch_bounds__(256) void lu_nopiv_diag_kernel(float *W, long ld, int c0, int w, float eps) {
  constexpr int NB = NP_LEAF;
  __shared__ float S[NB][NB + 1];
  __shared__ float s_inv;
  __shared__ int s_skip;
  const int t = threadIdx.x;
  for (int idx = t; idx < w * w; idx += 256) {
    const int r = idx / w, c = idx % w;
    S[r][c] = W[(long)(c0 + r) * ld + c0 + c];
  }
  __syncthreads();
  for (int k = 0; k < w; ++k) {
    if (t == 0) {
      const float p = S[k][k];
      s_skip = fabsf(p) < eps;
      s_inv = 1.0f / p;
    }
    __syncthreads();
    if (!s_skip) {
      const float inv = s_inv;
      for (int i = k + 1 + t; i < w; i += 256) S[i][k] *= inv;
      __syncthreads();
      const int m = w - k - 1;
      for (int idx = t; idx < m * m; idx += 256) {
        const int i = k + 1 + idx / m, c = k + 1 + idx % m;
        S[i][c] = fmaf(-S[i][k], S[k][c], S[i][c]);
      }
    }
    __syncthreads();
  }
  for (int idx = t; idx < w * w; idx += 256) {
    const int r = idx / w, c = idx % w;
    W[(long)(c0 + r) * ld + c0 + c] = S[r][c];
  }
}

static void nopiv_gemm_sub(hipStream_t s, long M, long N, long K, View<const float> A,
                           View<const float> B, View<float> C, Mul mul) {
  if (mul == Mul::Bf16x1 || mul == Mul::Bf16x3 || mul == Mul::Bf16x6)
    gemm_bf16x(s, M, N, K, -1.0f, A, B, C, (int)mul);
  else
    gemm<float>(s, M, N, K, -1.0f, A, B, 1.0f, C, false);
}

static void lu_nopiv_rec(hipStream_t s, int n, float *W, long ld, int c0, int w, float eps, Mul mul) {
  if (w <= NP_LEAF) {
    hipLaunchKernelGGL(lu_nopiv_diag_kernel, dim3(1), dim3(256), 0, s, W, ld, c0, w, eps);
    MG_HIP(hipGetLastError());
    const int m2 = n - c0 - w;
    if (m2 > 0) {
      // L21 = A21 U11^-1  <=>  U11^T L21^T = A21^T (U11^T lower, non-unit)
      View<const float> U11t{W + (long)c0 * ld + c0, 1, ld};
      View<float> A21t{W + (long)(c0 + w) * ld + c0, 1, ld};
      trsm_left<float>(s, true, false, w, m2, U11t, A21t, mul);
    }
    return;
  }
  int w1 = ((w / 2 + NP_LEAF - 1) / NP_LEAF) * NP_LEAF;
  if (w1 >= w) w1 = w - NP_LEAF;
  const int w2 = w - w1;
  // left block column (all rows c0..n)
  lu_nopiv_rec(s, n, W, ld, c0, w1, eps, mul);
  View<const float> L11{W + (long)c0 * ld + c0, ld, 1};
  View<float> A12{W + (long)c0 * ld + c0 + w1, ld, 1};
  trsm_left<float>(s, true, true, w1, w2, L11, A12, mul);
  const int m2 = n - c0 - w1;
  View<const float> L21{W + (long)(c0 + w1) * ld + c0, ld, 1};
  View<float> A22{W + (long)(c0 + w1) * ld + c0 + w1, ld, 1};
  nopiv_gemm_sub(s, m2, w2, w1, L21, cview(A12), A22, mul);
  lu_nopiv_rec(s, n, W, ld, c0 + w1, w2, eps, mul);
}

// Pivot-free LU of the n x n row-major fp32 matrix W (unit-lower L, U).
void getrf_nopiv_rowmajor(hipStream_t s, int n, float *W, long ld, float eps, Mul mul) {
  lu_nopiv_rec(s, n, W, ld, 0, n, eps, mul);
}

}  // namespace mg
