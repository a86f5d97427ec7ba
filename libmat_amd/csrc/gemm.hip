// FP32 / FP64 MFMA GEMM for gfx950.
//
// Serves mat_mul / mat_gemm (mat.h:1842-1850, 2787-2796) and every trailing
// update of the LU / Cholesky factorizations and triangular solves.
//
// Shape of the kernel:
//  * 256 threads = 4 wavefronts in a 2x2 grid; a workgroup owns a BM x BN
//    tile of C, each wavefront a (BM/2) x (BN/2) sub-tile held in MFMA
//    accumulators (fp32: v_mfma_f32_32x32x2f32, fp64: v_mfma_f64_16x16x4f64).
//  * K is walked in BK-deep slabs, double-buffered in LDS: the next slab's
//    global loads are issued before the current slab's MFMAs, then parked in
//    the other LDS buffer; one barrier per slab.
//  * Both operands live in LDS as [k][mn] (mn contiguous) so every MFMA
//    fragment read is a conflict-free 32-lane contiguous ds_read.  Global
//    tiles that are k-contiguous are transposed on the way in.
//  * Inside a slab, lane group g = lane / FM feeds k = g*(BK/G) + s at step
//    s; A and B use the same assignment, so the product is exact for any
//    assignment and each lane reads one scalar per fragment.
//  * Tiles are mapped XCD-aware: consecutive workgroup ids land on different
//    XCDs, so ids are regrouped so that each XCD's L2 sees a compact band of
//    C tiles (grouped along M) that shares A rows and B columns.
#include "ops.hpp"

namespace mg {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<float> {
  static constexpr int FM = 32;
  static constexpr int NACC = 16;
  using acc_t = f32x16;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // accumulator register r of lane -> row inside the 32x32 fragment
  __device__ static int row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
};
template <>
struct Mf<double> {
  static constexpr int FM = 16;
  static constexpr int NACC = 4;
  using acc_t = f64x4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // v_mfma_f64_16x16x4f64 interleaves rows across the four lane groups:
  // register r of lane l holds row (l/16) + 4r, column l%16
  __device__ static int row(int r, int lane) { return (lane >> 4) + 4 * r; }
};

template <typename T>
struct GemmP {
  const T *A;
  const T *B;
  T *C;
  long lda, ldb, ldc;
  int M, N, K;
  T alpha, beta;
  int tiles_m, tiles_n, tiles;
  int lower;
  int prio;  // wave priority (wave_prio of the launch stream)
  // tile count: all tiles, or for lower products with square tiles only the
  // tiles on or below the diagonal (the ids dealt to the XCDs are all work)
};

// Loads one BK x MNT operand slab from global memory into registers.
// MN=true: element (mn, k) at p[mn + k*ld]; MN=false: at p[k + mn*ld].
template <typename T, int MNT, int BK, bool MN, int VW>
struct SlabLoader {
  static constexpr int NT = 256;
  static constexpr int NV = MNT * BK / (NT * VW);
  static_assert(NV * NT * VW == MNT * BK, "slab must split evenly");
  T r[NV][VW];

  __device__ void load(const T *p, long ld, int mn0, int k0, int MNdim, int Kdim) {
    const int t = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int idx = t + NT * v;
      int mm, kk;
      if (MN) {
        kk = idx / (MNT / VW);
        mm = (idx % (MNT / VW)) * VW;
      } else {
        mm = idx / (BK / VW);
        kk = (idx % (BK / VW)) * VW;
      }
      const int gmn = mn0 + mm, gk = k0 + kk;
      const T *src = MN ? p + (long)gmn + (long)gk * ld : p + (long)gk + (long)gmn * ld;
      bool full = MN ? (gmn + VW <= MNdim && gk < Kdim) : (gmn < MNdim && gk + VW <= Kdim);
      if (VW > 1 && full) {
        if constexpr (VW == 4 && sizeof(T) == 4) {
          float4 x = *reinterpret_cast<const float4 *>(src);
          r[v][0] = x.x; r[v][1] = x.y; r[v][2] = x.z; r[v][3] = x.w;
        } else if constexpr (VW == 2 && sizeof(T) == 8) {
          double2 x = *reinterpret_cast<const double2 *>(src);
          r[v][0] = x.x; r[v][1] = x.y;
        }
      } else {
#pragma unroll
        for (int c = 0; c < VW; ++c) {
          bool ok = MN ? (gmn + c < MNdim && gk < Kdim) : (gmn < MNdim && gk + c < Kdim);
          r[v][c] = ok ? src[c] : T(0);
        }
      }
    }
  }

  // Store into LDS laid out as S[k][mn] with row stride LDS_LD.
  template <int LDS_LD>
  __device__ void store(T *S) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int idx = t + NT * v;
      if (MN) {
        const int kk = idx / (MNT / VW);
        const int mm = (idx % (MNT / VW)) * VW;
#pragma unroll
        for (int c = 0; c < VW; ++c) S[kk * LDS_LD + mm + c] = r[v][c];
      } else {
        const int mm = idx / (BK / VW);
        const int kk = (idx % (BK / VW)) * VW;
#pragma unroll
        for (int c = 0; c < VW; ++c) S[(kk + c) * LDS_LD + mm] = r[v][c];
      }
    }
  }
};

// Workgroup id -> (tile_m, tile_n) with XCD regrouping and GROUP_M bands.
__device__ inline bool map_tile(int bid, int grid, int tiles, int tiles_m, int tiles_n,
                                int &tm, int &tn) {
  constexpr int XCD = 8, GROUP_M = 8;
  int id = bid;
  if (grid % XCD == 0) id = (bid % XCD) * (grid / XCD) + bid / XCD;
  if (id >= tiles) return false;
  const int per_group = GROUP_M * tiles_n;
  const int g = id / per_group;
  const int first_m = g * GROUP_M;
  const int gsz = min(tiles_m - first_m, GROUP_M);
  const int in = id % per_group;
  tm = first_m + in % gsz;
  tn = in / gsz;
  return true;
}

// Compact lower enumeration: id -> (tm, tn), tn <= tm, row by row (the
// first tiles_n rows form a triangle, later rows hold tiles_n tiles each).
__device__ inline void lower_tile(int id, int tiles_n, int &tm, int &tn) {
  const int tri = tiles_n * (tiles_n + 1) / 2;
  if (id < tri) {
    int r = (int)((sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= id) ++r;
    while (r * (r + 1) / 2 > id) --r;
    tm = r;
    tn = id - r * (r + 1) / 2;
  } else {
    const int q = id - tri;
    tm = tiles_n + q / tiles_n;
    tn = q % tiles_n;
  }
}

// LDS slab buffers of a tile: two (the next slab is stored while this one is
// read) for the large tiles; one for the 64 x 64 tiles, which the
// factorizations' critical path runs beside the side stream's GEMMs: the
// next slab waits in registers anyway, one barrier more per slab keeps a
// workgroup's LDS at 35 KB instead of 70, so it fits on a compute unit
// beside two split-BF16 workgroups.  The sums are the same either way.
template <int BM, int BN>
constexpr int slab_bufs() {
  return BM == 64 && BN == 64 ? 1 : 2;
}

template <typename T, int BM, int BN, int BK, bool A_MN, bool B_MN, bool VEC>
__device__ __forceinline__ void gemm_tile(const GemmP<T> &p, int tm, int tn, T (*As)[BK * (BM + 16 / sizeof(T))],
                                          T (*Bs)[BK * (BN + 16 / sizeof(T))]);

// MAP 0: one tile per workgroup in the banded order; 1: one tile per
// workgroup, compact lower order; 2: compact lower order, persistent (grid a
// multiple of 8: XCD x = blockIdx % 8 owns the ids [x * per, (x + 1) * per),
// per = tiles / 8 rounded up, which its workgroups take in turn; the last
// slab's barrier retires every LDS reader before the next tile).  Separate
// instantiations, so that the one-tile kernels keep their registers.
template <typename T, int BM, int BN, int BK, bool A_MN, bool B_MN, bool VEC, int MAP>
__global__ __launch_bounds__(256) void gemm_kernel(GemmP<T> p) {
  constexpr int PAD = 16 / (int)sizeof(T);
  constexpr int XCD = 8;
  __shared__ T As[slab_bufs<BM, BN>()][BK * (BM + PAD)];
  __shared__ T Bs[slab_bufs<BM, BN>()][BK * (BN + PAD)];
  if (p.prio) __builtin_amdgcn_s_setprio(3);
  if constexpr (MAP == 0) {
    int tm, tn;
    if (!map_tile(blockIdx.x, gridDim.x, p.tiles, p.tiles_m, p.tiles_n, tm, tn)) return;
    gemm_tile<T, BM, BN, BK, A_MN, B_MN, VEC>(p, tm, tn, As, Bs);
  } else if constexpr (MAP == 1) {
    const int grid = gridDim.x, bid = blockIdx.x;
    const int id = grid % XCD == 0 ? (bid % XCD) * (grid / XCD) + bid / XCD : bid;
    if (id >= p.tiles) return;
    int tm, tn;
    lower_tile(id, p.tiles_n, tm, tn);
    gemm_tile<T, BM, BN, BK, A_MN, B_MN, VEC>(p, tm, tn, As, Bs);
  } else {
    const int x = blockIdx.x % XCD, j = blockIdx.x / XCD, wgs = gridDim.x / XCD;
    const int per = (p.tiles + XCD - 1) / XCD;
    for (int k = j; k < per; k += wgs) {
      const int id = x * per + k;
      if (id >= p.tiles) break;
      int tm, tn;
      lower_tile(id, p.tiles_n, tm, tn);
      gemm_tile<T, BM, BN, BK, A_MN, B_MN, VEC>(p, tm, tn, As, Bs);
    }
  }
}

template <typename T, int BM, int BN, int BK, bool A_MN, bool B_MN, bool VEC>
__device__ __forceinline__ void gemm_tile(const GemmP<T> &p, int tm, int tn, T (*As)[BK * (BM + 16 / sizeof(T))],
                                          T (*Bs)[BK * (BN + 16 / sizeof(T))]) {
  using F = Mf<T>;
  constexpr int FM = F::FM;
  constexpr int G = 64 / FM;
  constexpr int SK = BK / G;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int TM = WTM / FM, TN = WTN / FM;
  constexpr int VW = VEC ? 16 / (int)sizeof(T) : 1;
  constexpr int PAD = 16 / (int)sizeof(T);
  constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;

  const int bm0 = tm * BM, bn0 = tn * BN;
  // triangular C: skip tiles wholly outside the requested triangle
  if (p.lower == 1 && bm0 + BM - 1 < bn0) return;  // strictly above the diagonal
  if (p.lower == 2 && bm0 > bn0 + BN - 1) return;  // strictly below the diagonal

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int g = lane / FM, li = lane % FM;

  typename F::acc_t acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < F::NACC; ++r) acc[a][b][r] = T(0);

  SlabLoader<T, BM, BK, A_MN, VW> la;
  SlabLoader<T, BN, BK, B_MN, VW> lb;
  const int nk = (p.K + BK - 1) / BK;
  if (nk > 0) {
    la.load(p.A, p.lda, bm0, 0, p.M, p.K);
    lb.load(p.B, p.ldb, bn0, 0, p.N, p.K);
    la.template store<LDA_S>(As[0]);
    lb.template store<LDB_S>(Bs[0]);
  }
  __syncthreads();

  constexpr int NBUF = slab_bufs<BM, BN>();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = NBUF == 2 ? (kt & 1) : 0;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load(p.A, p.lda, bm0, (kt + 1) * BK, p.M, p.K);
      lb.load(p.B, p.ldb, bn0, (kt + 1) * BK, p.N, p.K);
    }
    const T *as = As[cur];
    const T *bs = Bs[cur];
#pragma unroll
    for (int s = 0; s < SK; ++s) {
      const int k = g * SK + s;
      T af[TM], bf[TN];
#pragma unroll
      for (int a = 0; a < TM; ++a) af[a] = as[k * LDA_S + wm * WTM + a * FM + li];
#pragma unroll
      for (int b = 0; b < TN; ++b) bf[b] = bs[k * LDB_S + wn * WTN + b * FM + li];
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = F::mma(af[a], bf[b], acc[a][b]);
    }
    if (more) {
      if constexpr (NBUF == 1) __syncthreads();  // every reader of the one buffer is done
      la.template store<LDA_S>(As[NBUF == 2 ? cur ^ 1 : 0]);
      lb.template store<LDB_S>(Bs[NBUF == 2 ? cur ^ 1 : 0]);
    }
    __syncthreads();
  }

  // Epilogue: C = alpha*acc + beta*C (C not read when beta == 0).
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int j = bn0 + wn * WTN + b * FM + li;
      if (j >= p.N) continue;
#pragma unroll
      for (int r = 0; r < F::NACC; ++r) {
        const int i = bm0 + wm * WTM + a * FM + F::row(r, lane);
        if (i >= p.M || (p.lower == 1 && i < j) || (p.lower == 2 && i > j)) continue;
        T *c = p.C + (long)i + (long)j * p.ldc;
        T v = p.alpha * acc[a][b][r];
        if (p.beta != T(0)) v += p.beta * *c;
        *c = v;
      }
    }
}

template <typename T, bool A_MN, bool B_MN, int BM, int BN, int BK>
static void launch_tiles(hipStream_t s, GemmP<T> &p, bool vec, int grid_cap) {
  p.tiles_m = ceil_div(p.M, BM);
  p.tiles_n = ceil_div(p.N, BN);
  const bool compact = p.lower == 1 && BM == BN;
  if (compact) {
    const long tm = p.tiles_m, tn = p.tiles_n, d = std::min(tm, tn);
    p.tiles = (int)(d * (d + 1) / 2 + std::max(0L, tm - tn) * tn);
  } else {
    p.tiles = p.tiles_m * p.tiles_n;
  }
  int grid = p.tiles;
  if (grid >= 8) grid = (grid + 7) / 8 * 8;
  // the persistent form exists for the large-tile lower products (SYRK)
  constexpr bool can_persist = BM == 128 && BN == 128;
  const bool persist = can_persist && compact && grid_cap >= 8 && grid > grid_cap;
  if (persist) grid = grid_cap / 8 * 8;
  p.prio = wave_prio(s);
  KScope ks(s, "gemm", (p.lower ? 1.0 : 2.0) * p.M * p.N * p.K);
#define MG_GEMM_LAUNCH(MAP_)                                                                        \
  {                                                                                                 \
    if (vec)                                                                                        \
      hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, A_MN, B_MN, true, MAP_>), dim3(grid), dim3(256), \
                         0, s, p);                                                                  \
    else                                                                                            \
      hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, A_MN, B_MN, false, MAP_>), dim3(grid),          \
                         dim3(256), 0, s, p);                                                       \
  }
  if constexpr (can_persist) {
    if (persist) MG_GEMM_LAUNCH(2) else if (compact) MG_GEMM_LAUNCH(1) else MG_GEMM_LAUNCH(0)
  } else {
    if (compact) MG_GEMM_LAUNCH(1) else MG_GEMM_LAUNCH(0)
  }
#undef MG_GEMM_LAUNCH
  MG_HIP(hipGetLastError());
}

// Tile shape by problem shape.  Large products: 128 x 128 tiles walking K
// in 16-deep slabs.  Products with fewer 128 x 128 tiles than the device
// has CUs (the skinny updates on the critical path of the factorizations)
// get one workgroup per CU at most, and every slab then costs a memory
// round trip that nothing hides; 64 x 64 tiles (four times the workgroups,
// two per CU) with 64-deep slabs (32 for f64, for LDS) take a quarter of the
// round trips.
template <typename T, bool A_MN, bool B_MN>
static void launch_mode(hipStream_t s, GemmP<T> &p, bool vec, int in_place, int grid_cap) {
  const long tiles128 = (long)ceil_div(p.M, 128) * ceil_div(p.N, 128);
  if (in_place) {
    // C is an operand (B := D B, or its transpose B^T := B^T D^T): safe only
    // when every workgroup reads exactly the part of the operand it
    // overwrites, i.e. one tile along M (C is B) or along N (C is A)
    MG_CHECK(in_place == 2 ? p.M <= 128 : p.N <= 128, "gemm: in-place product spans more than one tile");
    // the other dimension in 64-wide tiles and K in deep slabs: these
    // products are small and latency-bound (one round trip per slab)
    constexpr int BKI = sizeof(T) == 4 ? 64 : 32;
    if (in_place == 2) launch_tiles<T, A_MN, B_MN, 128, 64, BKI>(s, p, vec, 0);
    else launch_tiles<T, A_MN, B_MN, 64, 128, BKI>(s, p, vec, 0);
  } else if (tiles128 < 256 && p.K >= 32) {
    constexpr int BKS = sizeof(T) == 4 ? 64 : 32;
    launch_tiles<T, A_MN, B_MN, 64, 64, BKS>(s, p, vec, grid_cap);
  } else {
    launch_tiles<T, A_MN, B_MN, 128, 128, 16>(s, p, vec, grid_cap);
  }
}

template <typename T>
void gemm(hipStream_t s, long M, long N, long K, T alpha, View<const T> A, View<const T> B,
          T beta, View<T> C, int tri, int grid_cap) {
  if (M <= 0 || N <= 0) return;
  if (C.rs != 1) {
    // Row-major C: compute C^T = B^T A^T into a column-major view.
    MG_CHECK(C.cs == 1 && tri == 0, "gemm: unsupported C layout");
    gemm<T>(s, N, M, K, alpha, B.t(), A.t(), beta, View<T>{C.p, 1, C.rs}, false, grid_cap);
    return;
  }
  MG_CHECK(M < (1L << 31) && N < (1L << 31) && K < (1L << 31), "gemm: dimension too large");
  GemmP<T> p{};
  p.C = C.p;
  p.ldc = C.cs;
  p.M = (int)M;
  p.N = (int)N;
  p.K = (int)K;
  p.alpha = alpha;
  p.beta = beta;
  MG_CHECK(tri >= 0 && tri <= 2, "gemm: tri must be 0, 1 or 2");
  p.lower = tri;
  // A: element (i,k).  MN-contiguous when rs == 1.
  const bool a_mn = (A.rs == 1);
  p.A = A.p;
  p.lda = a_mn ? A.cs : A.rs;
  MG_CHECK(a_mn || A.cs == 1, "gemm: A needs a unit stride");
  // B: element (k,j).  K-contiguous when rs == 1, else contiguous along j.
  const bool b_mn = (B.rs != 1);
  p.B = B.p;
  p.ldb = b_mn ? B.rs : B.cs;
  MG_CHECK(!b_mn || B.cs == 1, "gemm: B needs a unit stride");
  constexpr int VW = 16 / sizeof(T);
  auto aligned = [&](const void *ptr, long ld) {
    return ((uintptr_t)ptr % 16 == 0) && (ld % VW == 0);
  };
  const bool vec = aligned(p.A, p.lda) && aligned(p.B, p.ldb);
  // C is the same buffer as an operand (B := D B, the in-place application
  // of an inverted diagonal block); disjoint sub-views of one matrix, as in
  // the factorizations' updates, are not in place
  const int in_place = (const void *)C.p == (const void *)B.p ? 2 : ((const void *)C.p == (const void *)A.p ? 1 : 0);
  if (a_mn && !b_mn) launch_mode<T, true, false>(s, p, vec, in_place, grid_cap);
  else if (a_mn && b_mn) launch_mode<T, true, true>(s, p, vec, in_place, grid_cap);
  else if (!a_mn && !b_mn) launch_mode<T, false, false>(s, p, vec, in_place, grid_cap);
  else launch_mode<T, false, true>(s, p, vec, in_place, grid_cap);
}

// ------------------------------------------------ folded tile rows ----
// Tp_r = D_r T(r, c0 : c0 + cols) for every block row r of a triangle in
// one launch (the folded tiles of the chained triangular solve, trsv.hip):
// block row r = blockIdx.y, tile blockIdx.x of its product.  Each block row
// is computed exactly as gemm() computes it on its own: C_r^T = T_r^T D_r^T
// into the column-major view of the row-major Tp_r, with the 64 x 64 tiles
// and deep slabs gemm() picks for such products (fewer 128 x 128 tiles than
// 256 and K >= 32); rows that gemm() would tile otherwise are left to it.
template <typename T>
struct FoldP {
  const T *Tm;
  long trs, tcs;  // the triangle's view: element (i, j) at Tm[i * trs + j * tcs]
  const T *D;     // inverted TB x TB diagonal blocks, row-major, in order
  T *Tp;
  int n, TB, lower, prio;
};

template <typename T>
__host__ __device__ inline bool fold_row_batched(int rrows, int cols) {
  return rrows >= 32 && (long)((cols + 127) / 128) * ((rrows + 127) / 128) < 256;
}

template <typename T, int BM, int BN, int BK, bool A_MN, bool VEC>
__global__ __launch_bounds__(256) void fold_rows_kernel(FoldP<T> f) {
  constexpr int PAD = 16 / (int)sizeof(T);
  __shared__ T As[slab_bufs<BM, BN>()][BK * (BM + PAD)];
  __shared__ T Bs[slab_bufs<BM, BN>()][BK * (BN + PAD)];
  const int r = blockIdx.y;
  const int ro = r * f.TB, rrows = min(f.TB, f.n - ro);
  const int c0 = f.lower ? 0 : ro + f.TB;
  const int cols = f.lower ? ro : f.n - c0;
  if (cols <= 0 || !fold_row_batched<T>(rrows, cols)) return;
  const int tiles_n = (rrows + BN - 1) / BN;
  const int id = blockIdx.x;
  if (id >= ((cols + BM - 1) / BM) * tiles_n) return;
  if (f.prio) __builtin_amdgcn_s_setprio(3);
  GemmP<T> p{};
  p.A = f.Tm + (long)ro * f.trs + (long)c0 * f.tcs;  // A'(i, k) = T(ro + k, c0 + i)
  p.lda = A_MN ? f.trs : f.tcs;
  p.B = f.D + (long)r * f.TB * f.TB;  // B'(k, j) = D_r(j, k)
  p.ldb = f.TB;
  p.C = f.Tp + fold_off(f.lower != 0, r, f.n, f.TB);
  p.ldc = fold_ld(f.lower != 0, r, f.n, f.TB);
  p.M = cols;
  p.N = rrows;
  p.K = rrows;
  p.alpha = T(1);
  p.beta = T(0);
  gemm_tile<T, BM, BN, BK, A_MN, false, VEC>(p, id / tiles_n, id % tiles_n, As, Bs);
}

template <typename T>
void gemm_fold_rows(hipStream_t s, bool lower, int n, int TB, View<const T> Tm, const T *D, T *Tp) {
  const int nb = ceil_div(n, TB);
  if (nb <= 1) return;
  const bool a_mn = Tm.cs == 1;
  MG_CHECK(a_mn || Tm.rs == 1, "gemm_fold_rows: the triangle needs a unit stride");
  constexpr int BM = 64, BN = 64, BK = sizeof(T) == 4 ? 64 : 32;
  constexpr int VW = 16 / (int)sizeof(T);
  const long lda = a_mn ? Tm.rs : Tm.cs;
  // every block row's operands start at multiples of TB (>= VW) elements
  const bool vec = (uintptr_t)Tm.p % 16 == 0 && lda % VW == 0 && (uintptr_t)D % 16 == 0 &&
                   TB % VW == 0;
  int maxt = 0;
  double work = 0.0;
  for (int r = 0; r < nb; ++r) {
    const int ro = r * TB, rrows = std::min(TB, n - ro);
    const int c0 = lower ? 0 : ro + TB;
    const int cols = lower ? ro : n - c0;
    if (cols <= 0) continue;
    if (!fold_row_batched<T>(rrows, cols)) {
      // a short last block row or a very wide one: gemm() tiles it differently
      View<const T> Dr{D + (long)r * TB * TB, TB, 1};
      View<T> C{Tp + fold_off(lower, r, n, TB), fold_ld(lower, r, n, TB), 1};
      gemm<T>(s, rrows, cols, rrows, T(1), Dr, Tm.sub(ro, c0), T(0), C, false);
      continue;
    }
    maxt = std::max(maxt, (int)(ceil_div(cols, BM) * ceil_div(rrows, BN)));
    work += 2.0 * rrows * cols * rrows;
  }
  if (maxt == 0) return;
  FoldP<T> f{Tm.p, Tm.rs, Tm.cs, D, Tp, n, TB, lower ? 1 : 0, wave_prio(s)};
  KScope ks(s, "gemm", work);
  const dim3 grid(maxt, nb);
  if (a_mn && vec) hipLaunchKernelGGL((fold_rows_kernel<T, BM, BN, BK, true, true>), grid, dim3(256), 0, s, f);
  else if (a_mn) hipLaunchKernelGGL((fold_rows_kernel<T, BM, BN, BK, true, false>), grid, dim3(256), 0, s, f);
  else if (vec) hipLaunchKernelGGL((fold_rows_kernel<T, BM, BN, BK, false, true>), grid, dim3(256), 0, s, f);
  else hipLaunchKernelGGL((fold_rows_kernel<T, BM, BN, BK, false, false>), grid, dim3(256), 0, s, f);
  MG_HIP(hipGetLastError());
}

template void gemm_fold_rows<float>(hipStream_t, bool, int, int, View<const float>, const float *, float *);
template void gemm_fold_rows<double>(hipStream_t, bool, int, int, View<const double>, const double *,
                                     double *);

template void gemm<float>(hipStream_t, long, long, long, float, View<const float>,
                          View<const float>, float, View<float>, int, int);
template void gemm<double>(hipStream_t, long, long, long, double, View<const double>,
                           View<const double>, double, View<double>, int, int);

}  // namespace mg
