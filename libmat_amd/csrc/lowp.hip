// Low-precision MFMA GEMMs for the mixed-precision solvers:
//   C(fp32) += alpha * A * B, A and B given in fp32.
//
// BF16 path (LU trailing updates): every fp32 operand is split exactly into
// bf16 terms x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), ...).
// passes = 1 multiplies x0*y0 only (bf16 accuracy, ~2^-8); passes = 3 adds
// x0*y1 + x1*y0 (~2^-16); passes = 6 adds x0*y2 + x1*y1 + x2*y0 (~fp32).
// Each bf16 x bf16 product is exact in fp32 and the MFMA accumulates in
// fp32, so the extra passes buy accuracy at 1/3 or 1/6 of the bf16 rate
// (still 5x / 2.6x the fp32 MFMA rate).
//
// FP8 path (Cholesky trailing updates): each operand is quantized to
// e4m3 with one power-of-two scale per 128 x 128 tile (128 rows/columns by
// 128 of K); the MFMA sums one 128-deep K block in fp32 and the epilogue of
// that block multiplies by the two tile scales before adding into the
// running fp32 accumulator.
//
// Both run on pre-packed operands: a pack kernel writes each operand as
// [mn][k] planes (k contiguous, zero padded to whole tiles), so the GEMM
// kernel moves 16-byte rows global -> LDS -> MFMA fragments with no
// bounds checks and no transposes.
#include <hip/hip_bf16.h>

#include "ops.hpp"

namespace mg {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int LP_BM = 128, LP_BN = 128;

__device__ inline unsigned short f32_to_bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// ---------------------------------------------------------------- pack --
// Src element (mn, k) = src[mn*rs + k*cs] (a strided view with one unit
// stride).  Dst plane p: dst[p*plane + mn*Kp + k], bf16 bits.
__global__ __launch_bounds__(256) void pack_bf16_kernel(const float *src, long rs, long cs, int MN,
                                                        int K, int Kp, int nplanes, long plane,
                                                        unsigned short *dst) {
  __shared__ float tile[32][33];
  const int m0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // read coalesced along the source's unit stride
  const bool mn_unit = (rs == 1);
  for (int yy = ty; yy < 32; yy += 8) {
    int mm, kk;
    if (mn_unit) {
      mm = m0 + tx;
      kk = k0 + yy;
    } else {
      kk = k0 + tx;
      mm = m0 + yy;
    }
    float v = 0.f;
    if (mm < MN && kk < K) v = src[(long)mm * rs + (long)kk * cs];
    if (mn_unit) tile[tx][yy] = v;  // tile[mn][k]
    else tile[yy][tx] = v;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int mm = m0 + yy, kk = k0 + tx;
    float x = tile[yy][tx];
    const long o = (long)mm * Kp + kk;
    for (int p = 0; p < nplanes; ++p) {
      const unsigned short h = f32_to_bf16_rne(x);
      dst[p * plane + o] = h;
      x = x - bf16_to_f32(h);
    }
  }
}

// FP8 e4m3 tile scales: one power of two per (128-row tile, 128-deep K
// block) so that max|x| / scale <= 448.
__global__ __launch_bounds__(256) void fp8_scale_kernel(const float *src, long rs, long cs, int MN,
                                                        int K, int ktiles, float *scale) {
  const int tm = blockIdx.y, tk = blockIdx.x;
  float mx = 0.f;
  for (int idx = threadIdx.x; idx < 128 * 128; idx += 256) {
    int mm, kk;
    if (rs == 1) {
      mm = tm * 128 + idx % 128;
      kk = tk * 128 + idx / 128;
    } else {
      kk = tk * 128 + idx % 128;
      mm = tm * 128 + idx / 128;
    }
    if (mm < MN && kk < K) mx = fmaxf(mx, fabsf(src[(long)mm * rs + (long)kk * cs]));
  }
  __shared__ float red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // smallest power of two s with max|x| / s <= 448
    const float m = red[0];
    float s = 1.f;
    if (m > 0.f && m < 3.0e38f) {
      int e;
      const float f = frexpf(m / 448.f, &e);  // m/448 = f * 2^e, f in [0.5, 1)
      s = ldexpf(1.f, f == 0.5f ? e - 1 : e);
    }
    scale[tm * ktiles + tk] = s;
  }
}

__device__ inline unsigned char f32_to_e4m3(float x) {
  // round-to-nearest-even into OCP e4m3 (bias 7, max 448, no infinities)
  const unsigned u = __float_as_uint(x);
  const unsigned sign = (u >> 24) & 0x80u;
  float a = fabsf(x);
  if (!(a == a)) return (unsigned char)(sign | 0x7f);
  if (a >= 448.f) return (unsigned char)(sign | 0x7e);
  if (a < 0.015625f) {  // below 2^-6: subnormals, step 2^-9 (q == 8 encodes 2^-6)
    const float q = rintf(a * 512.f);
    return (unsigned char)(sign | (unsigned)q);
  }
  int e;
  float m = frexpf(a, &e);  // a = m * 2^e, m in [0.5, 1)
  int E = e - 1 + 7;        // biased exponent of the 1.mmm form, >= 1 here
  float frac = m * 2.f - 1.f;  // [0,1)
  float q = rintf(frac * 8.f);
  if (q >= 8.f) {
    q = 0.f;
    E += 1;
  }
  if (E > 15 || (E == 15 && q > 6.f)) return (unsigned char)(sign | 0x7e);
  return (unsigned char)(sign | (E << 3) | (unsigned)q);
}

__global__ __launch_bounds__(256) void pack_fp8_kernel(const float *src, long rs, long cs, int MN,
                                                       int K, int Kp, int ktiles,
                                                       const float *scale, unsigned char *dst) {
  __shared__ float tile[32][33];
  const int m0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const bool mn_unit = (rs == 1);
  for (int yy = ty; yy < 32; yy += 8) {
    int mm, kk;
    if (mn_unit) {
      mm = m0 + tx;
      kk = k0 + yy;
    } else {
      kk = k0 + tx;
      mm = m0 + yy;
    }
    float v = 0.f;
    if (mm < MN && kk < K) v = src[(long)mm * rs + (long)kk * cs];
    if (mn_unit) tile[tx][yy] = v;
    else tile[yy][tx] = v;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int mm = m0 + yy, kk = k0 + tx;
    const float s = scale[(mm / 128) * ktiles + kk / 128];
    dst[(long)mm * Kp + kk] = f32_to_e4m3(tile[yy][tx] / s);
  }
}

// ---------------------------------------------------------------- GEMM --
struct LpP {
  const unsigned short *A;  // [planes][Mp][Kp]
  const unsigned short *B;  // [planes][Np][Kp]
  long planeA, planeB;
  float *C;
  long ldc;  // C column-major (rs = 1)
  int M, N, Kp;
  float alpha;
  int tiles_m, tiles_n, tiles;
  int lower;
  int vgrid;  // workgroup ids of the tile map (tiles rounded up to 8)
  // persistent grid (fewer workgroups than ids): per-XCD claim counters,
  // else null (one tile per workgroup)
  unsigned *ctr;
};

// Workgroup id -> (tm, tn).  The ids are regrouped by XCD; lower products
// (square tiles) enumerate only the tiles on or below the diagonal, row by
// row (the first tiles_n rows form a triangle, later rows hold tiles_n
// tiles each), so every XCD gets an equal share of real work; other
// products go in bands of GROUP_M tile rows.
__device__ inline void lp_tile_of(int id, int tiles_m, int tiles_n, int &tm, int &tn, int lower);
__device__ inline bool lp_map_tile(int bid, int grid, int tiles, int tiles_m, int tiles_n, int &tm,
                                   int &tn, int lower) {
  constexpr int XCD = 8;
  int id = bid;
  if (grid % XCD == 0) id = (bid % XCD) * (grid / XCD) + bid / XCD;
  if (id >= tiles) return false;
  lp_tile_of(id, tiles_m, tiles_n, tm, tn, lower);
  return true;
}
__device__ inline void lp_tile_of(int id, int tiles_m, int tiles_n, int &tm, int &tn, int lower) {
  constexpr int GROUP_M = 8;
  if (lower) {
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
    return;
  }
  const int per_group = GROUP_M * tiles_n;
  const int g = id / per_group;
  const int first_m = g * GROUP_M;
  const int gsz = min(tiles_m - first_m, GROUP_M);
  const int in = id % per_group;
  tm = first_m + in % gsz;
  tn = in / gsz;
}

// 256 threads, 2x2 wavefronts, 64x64 per wavefront (2x2 MFMA 32x32x16).
// NP planes per operand; PASSES products per K step (see file comment).
//
// Operands move global -> LDS with the direct load (global_load_lds, 16 bytes
// a lane, no registers in between) into one 32-deep slab buffer per
// workgroup (48 KB for three planes), so three workgroups share a compute
// unit and one's products cover the others' loads.  A buffer holds 128 rows
// of 32 bf16 (64 bytes), row r's 16-byte chunk c in slot c ^ h(r),
// h(r) = ((r >> 2) ^ (r >> 1)) & 3, so that the fragment reads (32 rows, one
// chunk) spread over all banks; a lane of the direct load writes slot
// lane % 4 of row lane / 4 of its wavefront's 16-row block and so loads
// chunk slot ^ h(row).
// The MFMA takes B's fragment as its row operand and A's as its column
// operand (it computes C^T), so in the accumulator a lane's 32 lanes run
// along C's rows: each epilogue access of column-major C is one 128-byte
// run per half-wavefront, and all of a lane's C values are read before any
// is written.
typedef __attribute__((address_space(3))) void lp_lds_void;
typedef const __attribute__((address_space(1))) void lp_glob_void;

__device__ __forceinline__ int lp_swz(int r) { return ((r >> 2) ^ (r >> 1)) & 3; }

template <int NP, int PASSES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void gemm_bf16x_kernel(LpP p) {
  constexpr int BM = LP_BM, BN = LP_BN, BK = 32;
  constexpr int TILE = BM * BK;  // bf16 per plane and operand (8 KB)
  __shared__ __align__(16) unsigned short sm[2][NP][TILE];  // [A, B][plane]
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int g = lane >> 5, li = lane & 31;
  // one tile per workgroup, or (a grid capped below the tile count, with a
  // multiple of 8 workgroups) tiles claimed one at a time: workgroup b takes
  // the next id of XCD b % 8's share of the tile map (so the map's XCD
  // grouping holds), and workgroups that start late, behind other kernels
  // on their compute units, simply take fewer.  The counter wraps to zero
  // on its last claim (every workgroup ends on exactly one claim past the
  // share), ready for the next launch.
  // (the claim goes to the lanes through the slab buffer, idle between
  // tiles: no LDS beyond the 48 KB of slabs)
  int *s_claim = reinterpret_cast<int *>(&sm[0][0][0]);
  const int xcd = blockIdx.x % 8, per = p.vgrid / 8;
  for (int vid = blockIdx.x, turn = 0;; vid += gridDim.x, ++turn) {
  int tm, tn;
  if (p.ctr) {
    __syncthreads();  // every lane is done with the slab buffer (and the previous claim)
    if (t == 0)
      *s_claim = (int)__builtin_amdgcn_atomic_inc32(p.ctr + xcd, (unsigned)(per + gridDim.x / 8 - 1),
                                                    __ATOMIC_RELAXED, "agent");
    __syncthreads();
    const int j = *s_claim;
    __syncthreads();  // read before the tile's first slab lands in the buffer
    if (j >= per) break;
    const int id = xcd * per + j;
    if (id >= p.tiles) continue;
    lp_tile_of(id, p.tiles_m, p.tiles_n, tm, tn, p.lower);
  } else {
    if (turn > 0) break;
    if (!lp_map_tile(vid, p.vgrid, p.tiles, p.tiles_m, p.tiles_n, tm, tn, p.lower)) continue;
  }
  const int bm0 = tm * BM, bn0 = tn * BN;
  if (p.lower && bm0 + BM - 1 < bn0) continue;

  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

  // direct loads: instruction q of wavefront w fills rows (4q + w) * 16 ..
  // + 15 of a buffer; this lane's row and the global chunk it fetches
  long goffA[2], goffB[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int r = (4 * q + wave) * 16 + (lane >> 2);
    const int c = (lane & 3) ^ lp_swz(r);
    goffA[q] = (long)(bm0 + r) * p.Kp + c * 8;
    goffB[q] = (long)(bn0 + r) * p.Kp + c * 8;
  }
  // fragment offsets: row (x * 32 + li) of the wavefront's 64-row block,
  // chunk 2 * ks + g
  const int hl = lp_swz(li);
  const int off0 = li * BK + ((0 + g) ^ hl) * 8, off1 = li * BK + ((2 + g) ^ hl) * 8;

  const int nk = p.Kp / BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int k0 = kt * BK;
#pragma unroll
    for (int pl = 0; pl < NP; ++pl)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int dst = (4 * q + wave) * 512;  // bf16 offset of the 1 KB block
        __builtin_amdgcn_global_load_lds((lp_glob_void *)(p.A + pl * p.planeA + goffA[q] + k0),
                                         (lp_lds_void *)&sm[0][pl][dst], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((lp_glob_void *)(p.B + pl * p.planeB + goffB[q] + k0),
                                         (lp_lds_void *)&sm[1][pl][dst], 16, 0, 0);
      }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this slab is in LDS
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK / 16; ++ks) {
      const int off = ks ? off1 : off0;
      bf16x8 fa[NP][2], fb[NP][2];
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          fa[pl][x] = *reinterpret_cast<const bf16x8 *>(&sm[0][pl][(wm * 64 + x * 32) * BK + off]);
          fb[pl][x] = *reinterpret_cast<const bf16x8 *>(&sm[1][pl][(wn * 64 + x * 32) * BK + off]);
        }
      // small terms first, leading term last; the four accumulators take
      // turns so that consecutive MFMAs are independent
#define LP_PROD(PA, PB)                                                                   \
  _Pragma("unroll") for (int a = 0; a < 2; ++a)                                           \
  _Pragma("unroll") for (int b = 0; b < 2; ++b)                                           \
      acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[PB][b], fa[PA][a], acc[a][b], 0, 0, 0);
      if constexpr (PASSES >= 6) {
        LP_PROD(0, 2 % NP)
        LP_PROD(1 % NP, 1 % NP)
        LP_PROD(2 % NP, 0)
      }
      if constexpr (PASSES >= 3) {
        LP_PROD(0, 1 % NP)
        LP_PROD(1 % NP, 0)
      }
      LP_PROD(0, 0)
#undef LP_PROD
    }
    __syncthreads();  // every wavefront is done with the slab before the next one lands
  }
  // acc[a][b][r] = C(i, j) with i = bm0 + wm*64 + a*32 + li (a lane's own
  // row) and j = bn0 + wn*64 + b*32 + 8*(r/4) + 4*g + r%4
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int i = bm0 + wm * 64 + a * 32 + li;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float *cb = p.C + (long)i + (long)(bn0 + wn * 64 + b * 32 + 4 * g) * p.ldc;
      const int jb = bn0 + wn * 64 + b * 32 + 4 * g;
      float cv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int jo = 8 * (r >> 2) + (r & 3);
        const bool in = i < p.M && jb + jo < p.N && !(p.lower && i < jb + jo);
        cv[r] = in ? cb[(long)jo * p.ldc] : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int jo = 8 * (r >> 2) + (r & 3);
        const bool in = i < p.M && jb + jo < p.N && !(p.lower && i < jb + jo);
        if (in) cb[(long)jo * p.ldc] = cv[r] + p.alpha * acc[a][b][r];
      }
    }
  }
  }
}

// FP8: A, B packed e4m3 [mn][Kp]; scales per (128-tile, 128-K block).
struct F8P {
  const unsigned char *A, *B;
  const float *sA, *sB;  // [tiles][ktiles]
  float *C;
  long ldc;
  int M, N, Kp, ktiles;
  float alpha;
  int tiles_m, tiles_n, tiles;
  int lower;
};

__global__ __launch_bounds__(256) void gemm_fp8_kernel(F8P p) {
  constexpr int BM = 128, BN = 128, BK = 128, LDK = BK + 16;
  __shared__ unsigned char As[BM * LDK];
  __shared__ unsigned char Bs[BN * LDK];
  int tm, tn;
  if (!lp_map_tile(blockIdx.x, gridDim.x, p.tiles, p.tiles_m, p.tiles_n, tm, tn, p.lower)) return;
  const int bm0 = tm * BM, bn0 = tn * BN;
  if (p.lower && bm0 + BM - 1 < bn0) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int g = lane >> 5, li = lane & 31;
  f32x16 acc[2][2], part[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
  const int nk = p.Kp / BK;
  for (int kt = 0; kt < nk; ++kt) {
    // 128 rows x 128 bytes per operand: 1024 16-byte chunks, 4 per thread
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = t + 256 * r, row = idx >> 3, ch = (idx & 7) * 16;
      *reinterpret_cast<uint4 *>(&As[row * LDK + ch]) =
          *reinterpret_cast<const uint4 *>(p.A + (long)(bm0 + row) * p.Kp + kt * BK + ch);
      *reinterpret_cast<uint4 *>(&Bs[row * LDK + ch]) =
          *reinterpret_cast<const uint4 *>(p.B + (long)(bn0 + row) * p.Kp + kt * BK + ch);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) part[a][b][r] = 0.f;
#pragma unroll
    for (int ks = 0; ks < BK / 64; ++ks) {
      i32x8 fa[2], fb[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        fa[x] = *reinterpret_cast<const i32x8 *>(&As[(wm * 64 + x * 32 + li) * LDK + ks * 64 + g * 32]);
        fb[x] = *reinterpret_cast<const i32x8 *>(&Bs[(wn * 64 + x * 32 + li) * LDK + ks * 64 + g * 32]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          part[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[a], fb[b], part[a][b], 0, 0,
                                                                       0, 127, 0, 127);
    }
    const float sc = p.sA[tm * p.ktiles + kt] * p.sB[tn * p.ktiles + kt];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[a][b][r] = fmaf(part[a][b][r], sc, acc[a][b][r]);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = bn0 + wn * 64 + b * 32 + li;
      if (j >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = bm0 + wm * 64 + a * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
        if (i >= p.M || (p.lower && i < j)) continue;
        float *c = p.C + (long)i + (long)j * p.ldc;
        *c += p.alpha * acc[a][b][r];
      }
    }
}

static long roundup(long x, long m) { return (x + m - 1) / m * m; }

// Normalizes the problem so that C is column-major; returns false if C is
// row-major and the caller must transpose the problem.
// Tiles launched: all, or (lower) those on or below the diagonal.
static int lp_tile_count(int tiles_m, int tiles_n, int lower) {
  if (!lower) return tiles_m * tiles_n;
  const long tm = tiles_m, tn = tiles_n, d = std::min(tm, tn);
  return (int)(d * (d + 1) / 2 + std::max(0L, tm - tn) * tn);
}

void gemm_bf16x(hipStream_t s, long M, long N, long K, float alpha, View<const float> A,
                View<const float> B, View<float> C, int passes, bool lower_only, int grid_cap) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  if (C.rs != 1) {
    MG_CHECK(C.cs == 1 && !lower_only, "gemm_bf16x: unsupported C layout");
    gemm_bf16x(s, N, M, K, alpha, B.t(), A.t(), View<float>{C.p, 1, C.rs}, passes, false,
               grid_cap);
    return;
  }
  MG_CHECK(passes == 1 || passes == 3 || passes == 6, "gemm_bf16x: passes must be 1, 3 or 6");
  const int np = passes == 1 ? 1 : (passes == 3 ? 2 : 3);
  const long Mp = roundup(M, LP_BM), Np = roundup(N, LP_BN), Kp = roundup(K, 32);
  Workspace &ws = workspace(s);
  unsigned short *Ap = (unsigned short *)ws.get("lp.A", (size_t)np * Mp * Kp * 2);
  unsigned short *Bp = (unsigned short *)ws.get("lp.B", (size_t)np * Np * Kp * 2);
  // A(i,k) = A.p[i*rs + k*cs] -> [mn=i][k];  B(k,j) = B.p[k*rs + j*cs] -> [mn=j][k]
  {
  KScope ks(s, "pack", (double)(M + N) * K * 4 + (double)(Mp + Np) * Kp * 2 * np);
  hipLaunchKernelGGL(pack_bf16_kernel, dim3(Kp / 32, Mp / 32), dim3(256), 0, s, A.p, A.rs, A.cs,
                     (int)M, (int)K, (int)Kp, np, Mp * Kp, Ap);
  hipLaunchKernelGGL(pack_bf16_kernel, dim3(Kp / 32, Np / 32), dim3(256), 0, s, B.p, B.cs, B.rs,
                     (int)N, (int)K, (int)Kp, np, Np * Kp, Bp);
  MG_HIP(hipGetLastError());
  }
  LpP p{};
  p.A = Ap;
  p.B = Bp;
  p.planeA = Mp * Kp;
  p.planeB = Np * Kp;
  p.C = C.p;
  p.ldc = C.cs;
  p.M = (int)M;
  p.N = (int)N;
  p.Kp = (int)Kp;
  p.alpha = alpha;
  p.tiles_m = (int)(Mp / LP_BM);
  p.tiles_n = (int)(Np / LP_BN);
  p.lower = lower_only ? 1 : 0;
  p.tiles = lp_tile_count(p.tiles_m, p.tiles_n, p.lower);
  int grid = p.tiles;
  if (grid >= 8) grid = (grid + 7) / 8 * 8;
  p.vgrid = grid;
  if (grid_cap >= 8 && grid > grid_cap) {
    grid = grid_cap / 8 * 8;
    bool fresh = false;
    p.ctr = (unsigned *)ws.get("lp.ctr", 8 * sizeof(unsigned), &fresh);
    if (fresh) fill_u32(s, p.ctr, 8, 0u);
  }
  KScope ks(s, "gemm_bf16x", 2.0 * passes * M * N * K);
  if (passes == 1) hipLaunchKernelGGL((gemm_bf16x_kernel<1, 1>), dim3(grid), dim3(256), 0, s, p);
  else if (passes == 3) hipLaunchKernelGGL((gemm_bf16x_kernel<2, 3>), dim3(grid), dim3(256), 0, s, p);
  else hipLaunchKernelGGL((gemm_bf16x_kernel<3, 6>), dim3(grid), dim3(256), 0, s, p);
  MG_HIP(hipGetLastError());
}

void gemm_fp8(hipStream_t s, long M, long N, long K, float alpha, View<const float> A,
              View<const float> B, View<float> C, bool lower_only) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  if (C.rs != 1) {
    MG_CHECK(C.cs == 1 && !lower_only, "gemm_fp8: unsupported C layout");
    gemm_fp8(s, N, M, K, alpha, B.t(), A.t(), View<float>{C.p, 1, C.rs}, false);
    return;
  }
  const long Mp = roundup(M, 128), Np = roundup(N, 128), Kp = roundup(K, 128);
  const int ktiles = (int)(Kp / 128);
  Workspace &ws = workspace(s);
  unsigned char *Ap = (unsigned char *)ws.get("f8.A", (size_t)Mp * Kp);
  unsigned char *Bp = (unsigned char *)ws.get("f8.B", (size_t)Np * Kp);
  float *sA = (float *)ws.get("f8.sA", (size_t)(Mp / 128) * ktiles * 4);
  float *sB = (float *)ws.get("f8.sB", (size_t)(Np / 128) * ktiles * 4);
  {
  KScope ks(s, "pack", (double)(M + N) * K * 8 + (double)(Mp + Np) * Kp);
  hipLaunchKernelGGL(fp8_scale_kernel, dim3(ktiles, Mp / 128), dim3(256), 0, s, A.p, A.rs, A.cs,
                     (int)M, (int)K, ktiles, sA);
  hipLaunchKernelGGL(fp8_scale_kernel, dim3(ktiles, Np / 128), dim3(256), 0, s, B.p, B.cs, B.rs,
                     (int)N, (int)K, ktiles, sB);
  hipLaunchKernelGGL(pack_fp8_kernel, dim3(Kp / 32, Mp / 32), dim3(256), 0, s, A.p, A.rs, A.cs,
                     (int)M, (int)K, (int)Kp, ktiles, sA, Ap);
  hipLaunchKernelGGL(pack_fp8_kernel, dim3(Kp / 32, Np / 32), dim3(256), 0, s, B.p, B.cs, B.rs,
                     (int)N, (int)K, (int)Kp, ktiles, sB, Bp);
  MG_HIP(hipGetLastError());
  }
  F8P p{};
  p.A = Ap;
  p.B = Bp;
  p.sA = sA;
  p.sB = sB;
  p.C = C.p;
  p.ldc = C.cs;
  p.M = (int)M;
  p.N = (int)N;
  p.Kp = (int)Kp;
  p.ktiles = ktiles;
  p.alpha = alpha;
  p.tiles_m = (int)(Mp / 128);
  p.tiles_n = (int)(Np / 128);
  p.lower = lower_only ? 1 : 0;
  p.tiles = lp_tile_count(p.tiles_m, p.tiles_n, p.lower);
  int grid = p.tiles;
  if (grid >= 8) grid = (grid + 7) / 8 * 8;
  KScope ks(s, "gemm_fp8", 2.0 * M * N * K);
  hipLaunchKernelGGL(gemm_fp8_kernel, dim3(grid), dim3(256), 0, s, p);
  MG_HIP(hipGetLastError());
}

}  // namespace mg
