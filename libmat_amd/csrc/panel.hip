// Partial-pivoting factorization of one leaf column block (the "panel") of
// the row-major working matrix, and the row interchanges it implies for the
// rest of the matrix.
//
// Reference semantics (mat__plu_blocked_, mat.h:4209-4247): for each column
// k of the block, pick the first row of max |a(i,k)|, i >= k (mat__iamax,
// mat.h:1247-1262), swap it with row k, and unless |pivot| < eps scale the
// column below by 1/pivot and apply the rank-1 update to the block's
// remaining columns.  Columns k >= n-1 are not pivoted (mat.h:4214).
//
// MI355X design: the panel's rows are spread over P workgroups, each
// holding its rows (all leaf columns) in registers for the whole leaf.  One
// column step costs one cross-workgroup rendezvous: every workgroup
// publishes its best candidate (|value|, row) plus the candidate row's
// values, and the owner of row k publishes row k; after the rendezvous every
// workgroup independently reduces the P candidates (same data, same answer),
// swaps, scales and updates its own rows.  With P == 1 the rendezvous is an
// LDS barrier.  The grid is launched cooperatively (all workgroups
// resident); every spin wait is bounded and reports a timeout instead of
// hanging.
#include <algorithm>

#include "ops.hpp"

namespace mg {

constexpr int PANEL_NT = 256;


constexpr long SPIN_LIMIT = 1L << 26;

template <typename T>
struct PanelP {
  T *W;
  long ld;
  int r0;      // first panel row == first panel column (diagonal block)
  int m;       // rows in the panel (r0 .. r0+m-1)
  int c0, w;   // panel columns
  int kmax;    // columns k >= kmax are not pivoted (n - 1)
  int nwg;     // workgroups taking part
  T eps;
  int *ipiv;
  unsigned *bar;      // rendezvous counter, zero at launch
  T *candv;           // [2][nwg]
  int *candi;         // [2][nwg]
  T *rowbuf;          // [2][nwg][WMAX]
  T *diagbuf;         // [2][WMAX]
  int *err;           // set to 1 on rendezvous timeout
};

template <typename T>
__device__ inline void better(T &bv, int &bi, T v, int i) {
  if (v > bv || (v == bv && i < bi)) {
    bv = v;
    bi = i;
  }
}

template <typename T, int WMAX, int RPT>
__global__ __launch_bounds__(PANEL_NT) void panel_gepp_kernel(PanelP<T> p) {
  constexpr int RPW = PANEL_NT * RPT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int b = blockIdx.x;
  const int P = p.nwg;
  const int rend = p.r0 + p.m;
  const int w = p.w;

  __shared__ T prow[WMAX], drow[WMAX];
  __shared__ T wv[PANEL_NT / 64];
  __shared__ int wi[PANEL_NT / 64];
  __shared__ int s_piv, s_wgw, s_abort;

  // Register rotation: at step j, slot 0 holds column j, slots 1..w-1-j the
  // columns still to factor, and the finished columns rotate to the tail.
  // Whole rows (all slots) move on a row interchange, so finished L entries
  // travel with their rows.  After w steps column c sits in slot
  // (c - w) mod WMAX.
  T a[RPT][WMAX];
  int row[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    row[q] = p.r0 + b * RPW + q * PANEL_NT + t;
    const T *src = p.W + (long)row[q] * p.ld + p.c0;
#pragma unroll
    for (int c = 0; c < WMAX; ++c) a[q][c] = (row[q] < rend && c < w) ? src[c] : T(0);
  }
  if (t == 0) s_abort = 0;

  int step = 0;
  for (int j = 0; j < w; ++j) {
    const int k = p.c0 + j;
    if (k < p.kmax) {
      // 1. candidate among this thread's rows (first row of max |a|)
      T bv = T(-1);
      int bi = 0x7fffffff;
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (row[q] < rend && row[q] >= k) better(bv, bi, (T)fabs(a[q][0]), row[q]);
      // 2. wavefront reduction
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        T ov = __shfl_xor(bv, off);
        int oi = __shfl_xor(bi, off);
        better(bv, bi, ov, oi);
      }
      if (lane == 0) {
        wv[wave] = bv;
        wi[wave] = bi;
      }
      __syncthreads();
      if (t == 0) {
#pragma unroll
        for (int w2 = 1; w2 < PANEL_NT / 64; ++w2) better(wv[0], wi[0], wv[w2], wi[w2]);
        if (wi[0] == 0x7fffffff) wi[0] = k;  // no comparable value (all NaN)
      }
      __syncthreads();
      const int wg_idx = wi[0];
      const int par = step & 1;

      if (P == 1) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] == wg_idx)
#pragma unroll
            for (int c = 0; c < WMAX; ++c) prow[c] = a[q][c];
          if (row[q] == k)
#pragma unroll
            for (int c = 0; c < WMAX; ++c) drow[c] = a[q][c];
        }
        if (t == 0) s_piv = wg_idx;
        __syncthreads();
      } else {
        // publish candidate, candidate row and (owner only) row k
        if (t == 0) {
          p.candv[par * P + b] = wv[0];
          p.candi[par * P + b] = wg_idx;
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] == wg_idx) {
            T *dst = p.rowbuf + ((long)par * P + b) * WMAX;
#pragma unroll
            for (int c = 0; c < WMAX; ++c) dst[c] = a[q][c];
          }
          if (row[q] == k) {
            T *dst = p.diagbuf + (long)par * WMAX;
#pragma unroll
            for (int c = 0; c < WMAX; ++c) dst[c] = a[q][c];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (t == 0) {
          __hip_atomic_fetch_add(p.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned target = (unsigned)P * (unsigned)(step + 1);
          long spins = 0;
          while (__hip_atomic_load(p.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > SPIN_LIMIT) {
              s_abort = 1;
              __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (s_abort) return;
        if (t < 64) {
          // every workgroup reduces the same P candidates
          T cv = T(-1);
          int ci = 0x7fffffff, cw = 0;
          for (int x = lane; x < P; x += 64) {
            T v = p.candv[par * P + x];
            int i = p.candi[par * P + x];
            if (v > cv || (v == cv && i < ci)) {
              cv = v;
              ci = i;
              cw = x;
            }
          }
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) {
            T ov = __shfl_xor(cv, off);
            int oi = __shfl_xor(ci, off);
            int ow = __shfl_xor(cw, off);
            if (ov > cv || (ov == cv && oi < ci)) {
              cv = ov;
              ci = oi;
              cw = ow;
            }
          }
          if (lane == 0) {
            s_piv = ci;
            s_wgw = cw;
          }
        }
        __syncthreads();
        if (t < WMAX) {
          const T dv = p.diagbuf[(long)par * WMAX + t];
          drow[t] = dv;
          prow[t] = (s_piv == k) ? dv : p.rowbuf[((long)par * P + s_wgw) * WMAX + t];
        }
        __syncthreads();
      }

      // 3. swap rows k and piv (all slots)
      const int piv = s_piv;
      if (b == 0 && t == 0) p.ipiv[k] = piv;
      if (piv != k) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] == k) {
#pragma unroll
            for (int c = 0; c < WMAX; ++c) a[q][c] = prow[c];
          } else if (row[q] == piv) {
#pragma unroll
            for (int c = 0; c < WMAX; ++c) a[q][c] = drow[c];
          }
        }
      }
      // 4. scale and rank-1 update of the columns still to factor
      //    (skipped for a near-zero pivot)
      const T pv = prow[0];
      if (!((T)fabs(pv) < p.eps)) {
        const T inv = T(1) / pv;
        const int rem = w - j;  // slots 1..rem-1 hold unfinished columns
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] > k && row[q] < rend) {
            const T l = a[q][0] * inv;
            a[q][0] = l;
#pragma unroll
            for (int c = 1; c < WMAX; ++c)
              if (c < rem) a[q][c] = fma(-prow[c], l, a[q][c]);
          }
        }
      }
      ++step;
      __syncthreads();  // prow/drow are rewritten next step
    }
    // 5. rotate: the finished column moves to the tail
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const T f = a[q][0];
#pragma unroll
      for (int c = 0; c + 1 < WMAX; ++c) a[q][c] = a[q][c + 1];
      a[q][WMAX - 1] = f;
    }
  }
  // after w rotations column c is in slot (c - w) mod WMAX
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (row[q] < rend) {
      T *dst = p.W + (long)row[q] * p.ld + p.c0;
#pragma unroll
      for (int sl = 0; sl < WMAX; ++sl) {
        int c = sl + w;
        if (c >= WMAX) c -= WMAX;
        if (c < w) dst[c] = a[q][sl];
      }
    }
  }
}

// Apply the row interchanges ipiv[k0 .. k0+w) (sequential swaps, global
// row indices) to every column of the row-major matrix outside the panel
// columns [c0, c0+w).  Wavefront 0 of each workgroup turns the swap
// sequence into an explicit (destination, source) list over the touched
// rows; then each thread moves its column's touched entries: all reads
// first, then all writes, so the permutation is applied in one pass.
template <typename T, int WMAX>
__global__ __launch_bounds__(256) void rowswap_kernel(T *W, long ld, int ncols, const int *ipiv,
                                                      int *perm, int k0, int w, int c0) {
  static_assert(WMAX <= 32, "window and out-of-window rows each use <= 32 lanes");
  __shared__ int dst[2 * WMAX], src[2 * WMAX];
  __shared__ int npair;
  const int t = threadIdx.x;
  if (t < 64) {
    // Lane i tracks window row k0+i (my_win = original row now there);
    // lane x also tracks the x-th distinct row below the window.
    const int lane = t;
    int my_win = k0 + lane, my_outr = -1, my_outc = -1, nout = 0;
    for (int kk = 0; kk < w; ++kk) {
      const int k = k0 + kk;
      const int pr = ipiv[k];
      if (pr == k) continue;
      const int a = __shfl(my_win, kk);
      int bsrc;
      if (pr < k0 + w) {
        bsrc = __shfl(my_win, pr - k0);
        if (lane == pr - k0) my_win = a;
      } else {
        const unsigned long long m = __ballot(my_outr == pr);
        int slot;
        if (m) {
          slot = __builtin_ctzll(m);
        } else {
          slot = nout++;
          if (lane == slot) my_outr = my_outc = pr;
        }
        bsrc = __shfl(my_outc, slot);
        if (lane == slot) my_outc = a;
      }
      if (lane == kk) my_win = bsrc;
    }
    const bool wpair = lane < w && my_win != k0 + lane;
    const bool opair = lane < nout && my_outc != my_outr;
    const unsigned long long mw = __ballot(wpair), mo = __ballot(opair);
    const unsigned long long below = (1ull << lane) - 1ull;
    const int nw = __popcll(mw);
    if (wpair) {
      const int at = __popcll(mw & below);
      dst[at] = k0 + lane;
      src[at] = my_win;
    }
    if (opair) {
      const int at = nw + __popcll(mo & below);
      dst[at] = my_outr;
      src[at] = my_outc;
    }
    if (lane == 0) npair = nw + __popcll(mo);
  }
  __syncthreads();
  const int np = npair;
  if (np == 0) return;
  if (blockIdx.x == 0 && t < 64) {
    // the row-index vector moves with the rows (libmat's Perm)
    int pv = 0;
    if (t < np) pv = perm[src[t]];
    __builtin_amdgcn_wave_barrier();
    if (t < np) perm[dst[t]] = pv;
  }
  int c = blockIdx.x * 256 + t;
  if (c >= c0) c += w;  // skip the panel's own columns
  if (c >= ncols) return;
  T v[2 * WMAX];
#pragma unroll
  for (int x = 0; x < 2 * WMAX; ++x)
    if (x < np) v[x] = W[(long)src[x] * ld + c];
#pragma unroll
  for (int x = 0; x < 2 * WMAX; ++x)
    if (x < np) W[(long)dst[x] * ld + c] = v[x];
}

template <typename T, int WMAX, int RPT>
static void launch_panel(hipStream_t s, PanelP<T> &p) {
  constexpr int RPW = PANEL_NT * RPT;
  p.nwg = ceil_div(p.m, RPW);
  void *args[] = {&p};
  if (p.nwg == 1) {
    hipLaunchKernelGGL((panel_gepp_kernel<T, WMAX, RPT>), dim3(1), dim3(PANEL_NT), 0, s, p);
  } else {
    MG_HIP(hipLaunchCooperativeKernel((const void *)panel_gepp_kernel<T, WMAX, RPT>,
                                      dim3(p.nwg), dim3(PANEL_NT), args, 0, s));
  }
  MG_HIP(hipGetLastError());
}

// Factor the leaf block W[c0:n, c0:c0+w] (w <= 32) and apply its row swaps
// to all other columns of the n x n row-major matrix.
template <typename T>
void lu_leaf(hipStream_t s, int n, T *W, long ld, int c0, int w, int *ipiv, int *perm, T eps) {
  constexpr int WMAX = 32;
  MG_CHECK(w <= WMAX, "lu_leaf: leaf too wide");
  Workspace &ws = workspace(s);
  const int m = n - c0;
  PanelP<T> p{};
  p.W = W;
  p.ld = ld;
  p.r0 = c0;
  p.m = m;
  p.c0 = c0;
  p.w = w;
  p.kmax = n - 1;
  p.eps = eps;
  p.ipiv = ipiv;
  const int maxwg = 1024;
  p.candv = (T *)ws.get("panel.candv", 2 * maxwg * sizeof(T));
  p.candi = (int *)ws.get("panel.candi", 2 * maxwg * sizeof(int));
  p.rowbuf = (T *)ws.get("panel.rowbuf", 2L * maxwg * WMAX * sizeof(T));
  p.diagbuf = (T *)ws.get("panel.diagbuf", 2 * WMAX * sizeof(T));
  p.err = (int *)ws.get("panel.err", sizeof(int));
  p.bar = (unsigned *)ws.get("panel.bar", sizeof(unsigned));
  // rows per workgroup: keep the number of workgroups in the rendezvous small
  // double rows use twice the registers: at most 2 rows per thread
  const int rpt_max = sizeof(T) == 4 ? 4 : 2;
  const int rpt = std::min(rpt_max, (m > 256 * 8) ? 4 : (m > 256 * 2 ? 2 : 1));
  const int nwg = ceil_div(m, 256 * rpt);
  MG_CHECK(nwg <= maxwg, "lu_leaf: matrix too tall");
  if (nwg > 1) fill_u32(s, p.bar, 1, 0u);
  if (rpt == 4) {
    if constexpr (sizeof(T) == 4) launch_panel<T, WMAX, 4>(s, p);
  } else if (rpt == 2) {
    launch_panel<T, WMAX, 2>(s, p);
  } else {
    launch_panel<T, WMAX, 1>(s, p);
  }
  const int other = n - w;
  hipLaunchKernelGGL((rowswap_kernel<T, WMAX>), dim3(other > 0 ? ceil_div(other, 256) : 1),
                     dim3(256), 0, s, W, ld, n, ipiv, perm, c0, w, c0);
  MG_HIP(hipGetLastError());
}

template void lu_leaf<float>(hipStream_t, int, float *, long, int, int, int *, int *, float);
template void lu_leaf<double>(hipStream_t, int, double *, long, int, int, int *, int *, double);

}  // namespace mg
