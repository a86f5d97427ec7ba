// Householder QR of a column-major m x n matrix (m >= n), in place on R,
// with the orthogonal factor Q accumulated on request.
//
// Reference semantics (mat_qr, mat.h:3907-4063; mat_qr_r, mat.h:4066-4192):
// reflector j of column x = R(j:, j) is v = [1, x(1:)/(x0 - beta)],
// beta = -sign(x0) ||x||, tau = -(x0 - beta)/beta (mat_householder,
// mat.h:3634-3716); a column with ||x|| < eps gets v = 0 and tau = 0 and is
// left as it is.  Every column c >= j of the block is updated as
// R(j:, c) -= tau (v . R(j:, c)) v, so R keeps the reflectors' rounding
// residue below its diagonal, as the reference leaves it.  Blocks of
// reflectors reach the trailing columns and Q in compact-WY form,
// H_0 ... H_{k-1} = I - Y T Y^T (mat__qr_build_T_, mat.h:3772-3827;
// applies, mat.h:3846-3903).  The reference switches to the unblocked
// algorithm below 64 columns; it is the same product of reflectors, so one
// blocked path serves every size here.
//
// MI355X design.  Blocks of 32 reflectors (the reference uses 16; fewer,
// deeper GEMMs).  One workgroup of 1024 threads factors a block's panel
// column by column out of L2 (norm, reflector, then the dot products of
// the block's remaining columns, one wavefront each, and their update); a
// one-wavefront kernel builds T from Y^T Y; the trailing columns of R and
// the columns of Q are updated by the library's MFMA GEMM (three products
// each).  Look-ahead: the caller's stream applies a block only to the next
// block's columns and factors that panel, while the library's side stream
// applies it to the rest of R and to Q, so the wide updates overlap the
// panel, which occupies one CU.
#include <mutex>

#include "ops.hpp"

namespace mg {

namespace {

constexpr int QR_NB = 32;   // reflectors per block
constexpr int QR_NT = 1024; // threads of the panel workgroup
constexpr int QR_NW = QR_NT / 64;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Reflectors jb .. jb+kb-1 of R (column-major, ld ldr) into Y (len x kb,
// len = m - jb, ld ldy; zeros above each reflector's unit entry) and tau;
// the block's own columns of R are updated in place.
template <typename T>
__global__ __launch_bounds__(QR_NT) void qr_panel_kernel(int m, int jb, int kb, T *R, long ldr,
                                                         T *Y, long ldy, T *tau_out, T eps) {
  __shared__ T red[QR_NW];
  __shared__ T wdot[QR_NB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int j = 0; j < kb; ++j) {
    const int col = jb + j;
    const long vlen = (long)m - col;
    const T *x = R + (long)col * ldr + col;
    T *y = Y + (long)j * ldy;
    // ||x||^2: every thread sums the wavefront partials in the same order,
    // so every thread holds the same norm and takes the same branch
    T sq[4] = {0, 0, 0, 0};
    long i = t;
    for (; i + 3L * QR_NT < vlen; i += 4L * QR_NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) sq[r] += x[i + (long)r * QR_NT] * x[i + (long)r * QR_NT];
    }
    for (; i < vlen; i += QR_NT) sq[0] += x[i] * x[i];
    T s = (sq[0] + sq[1]) + (sq[2] + sq[3]);
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    T nsq = 0;
#pragma unroll
    for (int q = 0; q < QR_NW; ++q) nsq += red[q];
    const T norm = sqrt(nsq);
    for (long i = t; i < j; i += QR_NT) y[i] = 0;
    T tau = 0;
    if (norm < eps) {
      for (long i = t; i < vlen; i += QR_NT) y[j + i] = 0;
    } else {
      const T x0 = x[0];
      const T beta = x0 >= 0 ? -norm : norm;
      const T v0 = x0 - beta;
      const T v0_inv = T(1) / v0;
      tau = -v0 / beta;
      for (long i = t; i < vlen; i += QR_NT) y[j + i] = i == 0 ? T(1) : x[i] * v0_inv;
    }
    if (t == 0) tau_out[j] = tau;
    // v complete; every thread has read x0 before column col changes
    __syncthreads();
    if (tau != 0) {
      // w_c = (v . R(col:, c)) tau for the block's columns c >= j, one
      // wavefront per column, eight partial sums (eight loads in flight)
      for (int c = wave; c < kb; c += QR_NW) {
        if (c < j) continue;
        const T *rc = R + (long)(jb + c) * ldr + col;
        const T *yv = y + j;
        T d[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) d[u] = 0;
        long i = lane;
        for (; i + 7 * 64 < vlen; i += 8 * 64) {
#pragma unroll
          for (int u = 0; u < 8; ++u) d[u] += yv[i + u * 64] * rc[i + u * 64];
        }
        for (; i < vlen; i += 64) d[0] += yv[i] * rc[i];
        T dd = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7]));
        dd = wave_sum(dd);
        if (lane == 0) wdot[c] = dd * tau;
      }
      __syncthreads();
      // R(col:, c) += (-w_c) v, in tiles of RG rows (RG per thread, QR_NT
      // apart) by eight columns: a tile's values are all loaded before any
      // is stored, one memory round trip per tile
      constexpr int RG = sizeof(T) == 4 ? 4 : 2;
      for (long i0 = t; i0 < vlen; i0 += (long)RG * QR_NT) {
        T vi[RG];
        bool ok[RG];
#pragma unroll
        for (int r = 0; r < RG; ++r) {
          ok[r] = i0 + (long)r * QR_NT < vlen;
          vi[r] = ok[r] ? y[j + i0 + (long)r * QR_NT] : T(0);
        }
        T *p = R + (long)(jb + j) * ldr + col + i0;
        for (int c = j; c < kb; c += 8) {
          T val[RG][8];
#pragma unroll
          for (int r = 0; r < RG; ++r)
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (ok[r] && c + u < kb) val[r][u] = p[(long)u * ldr + (long)r * QR_NT];
#pragma unroll
          for (int r = 0; r < RG; ++r)
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (ok[r] && c + u < kb)
                p[(long)u * ldr + (long)r * QR_NT] = val[r][u] + -wdot[c + u] * vi[r];
          p += 8 * ldr;
        }
      }
      __syncthreads();
    }
  }
}

// T (kb x kb, upper, ld kb) of the compact-WY form from G = Y^T Y (ld kb)
// and tau: T(j,j) = tau_j, T(i,j) = -tau_j sum_{p=i}^{j-1} T(i,p) G(p,j)
// (mat.h:3810-3823).  G and tau are staged in LDS first (all loads in
// flight at once); lane i then builds row i, which only reads its own row.
template <typename T>
__global__ __launch_bounds__(64) void qr_tfactor_kernel(int kb, const T *G, const T *tau, T *Tm) {
  __shared__ T Ts[QR_NB][QR_NB + 1];
  __shared__ T Gs[QR_NB][QR_NB + 1];
  __shared__ T taus[QR_NB];
  const int i = threadIdx.x;
  for (int x = i; x < kb * kb; x += 64) Gs[x % kb][x / kb] = G[x];
  if (i < kb) taus[i] = tau[i];
  __syncthreads();
  if (i >= kb) return;
  for (int c = 0; c < kb; ++c) Ts[i][c] = 0;
  for (int j = 0; j < kb; ++j) {
    if (i < j) {
      T sum = 0;
      for (int p = i; p < j; ++p) sum += Ts[i][p] * Gs[p][j];
      Ts[i][j] = -taus[j] * sum;
    } else if (i == j) {
      Ts[i][j] = taus[j];
    }
  }
  for (int c = 0; c < kb; ++c) Tm[i + (long)c * kb] = Ts[i][c];
}

// C (kb x N, ld ldc) = Y^T B for Y (len x kb, ld ldy) and B (len x N, ld
// ldb), kb, N <= 32, K = len long: split over chunks of QR_CH rows (one
// workgroup each, both tiles staged in LDS) into partials P[chunk][32*32],
// then summed in chunk order (deterministic, no atomics).  A GEMM tile
// would walk all of K alone.
constexpr int QR_CH = 64;
template <typename T>
__global__ __launch_bounds__(256) void qr_ytb_partial_kernel(long len, int kb, int N, const T *Y,
                                                             long ldy, const T *B, long ldb, T *P) {
  __shared__ T Ys[QR_CH][QR_NB + 1];
  __shared__ T Bs[QR_CH][QR_NB + 1];
  const long r0 = (long)blockIdx.x * QR_CH;
  const int t = threadIdx.x;
  for (int x = t; x < QR_CH * kb; x += 256) {
    const int r = x % QR_CH, c = x / QR_CH;
    Ys[r][c] = r0 + r < len ? Y[r0 + r + (long)c * ldy] : T(0);
  }
  for (int x = t; x < QR_CH * N; x += 256) {
    const int r = x % QR_CH, c = x / QR_CH;
    Bs[r][c] = r0 + r < len ? B[r0 + r + (long)c * ldb] : T(0);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < (QR_NB * QR_NB) / 256; ++q) {
    const int idx = t + 256 * q, i = idx % QR_NB, j = idx / QR_NB;
    if (i < kb && j < N) {
      T acc = 0;
#pragma unroll 8
      for (int r = 0; r < QR_CH; ++r) acc += Ys[r][i] * Bs[r][j];
      P[(long)blockIdx.x * (QR_NB * QR_NB) + idx] = acc;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void qr_ytb_reduce_kernel(int nch, int kb, int N, const T *P,
                                                            T *C, long ldc) {
  const int idx = blockIdx.x * 256 + threadIdx.x, i = idx % QR_NB, j = idx / QR_NB;
  if (idx >= QR_NB * QR_NB || i >= kb || j >= N) return;
  T acc = 0;
  for (int c = 0; c < nch; ++c) acc += P[(long)c * (QR_NB * QR_NB) + idx];
  C[i + (long)j * ldc] = acc;
}

template <typename T>
__global__ void qr_eye_kernel(long m, T *Q) {
  const long total = m * m;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x)
    Q[idx] = (idx % m == idx / m) ? T(1) : T(0);
}

}  // namespace

// Schedule (k = block of 32 reflectors).  Caller's stream: panel k, then
// T_k, then block k applied to block k+1's columns only, then panel k+1.
// Side stream, in order: block k applied to every column right of block
// k+1, then to Q.  The two never touch the same columns at once; a block's
// columns receive the blocks before it in order (the caller's update of
// block k+1's columns waits for the side stream's update of block k-1);
// Y and T alternate between two slots, and panel k+2 waits until the side
// stream is done with block k's slot.
template <typename T>
void geqrf(hipStream_t s, int m, int n, T *R, long ldr, T *Q, T eps) {
  MG_CHECK(m >= n && n >= 1, "geqrf: needs m >= n >= 1");
  MG_CHECK(ldr >= m, "geqrf: leading dimension below m");
  Workspace &ws = workspace(s);
  T *Ys[2] = {(T *)ws.get("qr.Y0", (size_t)m * QR_NB * sizeof(T)),
              (T *)ws.get("qr.Y1", (size_t)m * QR_NB * sizeof(T))};
  T *Ts[2] = {(T *)ws.get("qr.T0", QR_NB * QR_NB * sizeof(T)),
              (T *)ws.get("qr.T1", QR_NB * QR_NB * sizeof(T))};
  T *tau = (T *)ws.get("qr.tau", QR_NB * sizeof(T));
  T *G = (T *)ws.get("qr.G", QR_NB * QR_NB * sizeof(T));
  T *Cn = (T *)ws.get("qr.Cn", QR_NB * QR_NB * sizeof(T));    // caller's stream
  T *Cn2 = (T *)ws.get("qr.Cn2", QR_NB * QR_NB * sizeof(T));
  T *Ck = (T *)ws.get("qr.Ck", (size_t)QR_NB * n * sizeof(T));  // side stream
  T *Ck2 = (T *)ws.get("qr.Ck2", (size_t)QR_NB * n * sizeof(T));
  T *Cm = Q ? (T *)ws.get("qr.Cm", (size_t)m * QR_NB * sizeof(T)) : nullptr;
  T *Cm2 = Q ? (T *)ws.get("qr.Cm2", (size_t)m * QR_NB * sizeof(T)) : nullptr;
  const int nblk = (n + QR_NB - 1) / QR_NB;
  T *P = (T *)ws.get("qr.P", ((size_t)m / QR_CH + 1) * QR_NB * QR_NB * sizeof(T));
  // C (kb x N, ld kb) = Y^T B with B len x N (ld ldb), N <= QR_NB
  auto ytb = [&](const T *Yp, long len, int kb, const T *B, long ldb, int N, T *Cp) {
    const int nch = (int)((len + QR_CH - 1) / QR_CH);
    hipLaunchKernelGGL((qr_ytb_partial_kernel<T>), dim3(nch), dim3(256), 0, s, len, kb, N, Yp, len,
                       B, ldb, P);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL((qr_ytb_reduce_kernel<T>), dim3((QR_NB * QR_NB + 255) / 256), dim3(256), 0,
                       s, nch, kb, N, P, Cp, (long)kb);
    MG_HIP(hipGetLastError());
  };
  // events: 3k = Y_k and T_k ready; 3k+1 = side stream done with block k's
  // columns; 3k+2 = side stream done with block k (its slot is free)
  std::lock_guard<std::mutex> lk(lookahead_mutex());
  hipStream_t u = side_stream();
  hipEvent_t start = pooled_event(3 * nblk);
  MG_HIP(hipEventRecord(start, s));
  MG_HIP(hipStreamWaitEvent(u, start, 0));
  if (Q) {
    const long total = (long)m * m;
    const unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL((qr_eye_kernel<T>), dim3(blocks), dim3(256), 0, u, (long)m, Q);
    MG_HIP(hipGetLastError());
  }
  // R(jb:, c0:c0+w) -= Y (T^T (Y^T R(jb:, c0:c0+w)))
  auto apply_r = [&](hipStream_t st, const View<const T> &y, const View<const T> &tm, int kb,
                     long len, int jb, int c0, int w, T *C1, T *C2) {
    T *Rs = R + (long)c0 * ldr + jb;
    gemm<T>(st, kb, w, len, T(1), y.t(), View<const T>{Rs, 1, ldr}, T(0), View<T>{C1, 1, kb});
    gemm<T>(st, kb, w, kb, T(1), tm.t(), View<const T>{C1, 1, kb}, T(0), View<T>{C2, 1, kb});
    gemm<T>(st, len, w, kb, T(-1), y, View<const T>{C2, 1, kb}, T(1), View<T>{Rs, 1, ldr});
  };
  for (int k = 0; k < nblk; ++k) {
    const int jb = k * QR_NB;
    const int kb = std::min(QR_NB, n - jb);
    const long len = (long)m - jb;
    T *Y = Ys[k & 1];
    T *Tm = Ts[k & 1];
    if (k >= 2) MG_HIP(hipStreamWaitEvent(s, pooled_event(3 * (k - 2) + 2), 0));
    {
      KScope ks(s, "qr_panel", 2.0 * len * kb * sizeof(T));
      hipLaunchKernelGGL((qr_panel_kernel<T>), dim3(1), dim3(QR_NT), 0, s, m, jb, kb, R, ldr, Y,
                         len, tau, eps);
      MG_HIP(hipGetLastError());
    }
    const View<const T> y{Y, 1, len};
    // G = Y^T Y, then T
    ytb(Y, len, kb, Y, len, kb, G);
    hipLaunchKernelGGL((qr_tfactor_kernel<T>), dim3(1), dim3(64), 0, s, kb, G, tau, Tm);
    MG_HIP(hipGetLastError());
    const View<const T> tm{Tm, 1, kb};
    MG_HIP(hipEventRecord(pooled_event(3 * k), s));
    const int c1 = jb + kb;                      // block k+1's first column
    const int w1 = std::min(QR_NB, n - c1);      // its width (<= 0: none)
    const int c2 = c1 + std::max(w1, 0);         // the rest
    // side stream: block k on the columns right of block k+1, then on Q
    MG_HIP(hipStreamWaitEvent(u, pooled_event(3 * k), 0));
    if (c2 < n) apply_r(u, y, tm, kb, len, jb, c2, n - c2, Ck, Ck2);
    MG_HIP(hipEventRecord(pooled_event(3 * k + 1), u));
    if (Q) {
      // Q(:, jb:) -= ((Q(:, jb:) Y) T) Y^T
      T *Qt = Q + (long)jb * m;
      gemm<T>(u, m, kb, len, T(1), View<const T>{Qt, 1, m}, y, T(0), View<T>{Cm, 1, m});
      gemm<T>(u, m, kb, kb, T(1), View<const T>{Cm, 1, m}, tm, T(0), View<T>{Cm2, 1, m});
      gemm<T>(u, m, len, kb, T(-1), View<const T>{Cm2, 1, m}, y.t(), T(1), View<T>{Qt, 1, m});
    }
    MG_HIP(hipEventRecord(pooled_event(3 * k + 2), u));
    // caller's stream: block k on block k+1's columns, after block k-1's
    // update of them (side stream) is done
    if (w1 > 0) {
      if (k >= 1) MG_HIP(hipStreamWaitEvent(s, pooled_event(3 * (k - 1) + 1), 0));
      T *Rn = R + (long)c1 * ldr + jb;
      ytb(Y, len, kb, Rn, ldr, w1, Cn);
      gemm<T>(s, kb, w1, kb, T(1), tm.t(), View<const T>{Cn, 1, kb}, T(0), View<T>{Cn2, 1, kb});
      gemm<T>(s, len, w1, kb, T(-1), y, View<const T>{Cn2, 1, kb}, T(1), View<T>{Rn, 1, ldr});
    }
  }
  hipEvent_t done = pooled_event(3 * nblk + 1);
  MG_HIP(hipEventRecord(done, u));
  MG_HIP(hipStreamWaitEvent(s, done, 0));
}

template void geqrf<float>(hipStream_t, int, int, float *, long, float *, float);
template void geqrf<double>(hipStream_t, int, int, double *, long, double *, double);

}  // namespace mg
