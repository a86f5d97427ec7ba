// Tournament pivoting for one leaf column block of the row-major working
// matrix (fp32): the communication-avoiding LU panel of Grigori, Demmel and
// Xiang ("CALU: a communication optimal LU factorization algorithm", SIAM
// J. Matrix Anal. Appl. 32(4), 2011), offered to the mixed-precision solver
// as an alternative to libmat's partial pivoting (mat.h:4209-4247).  The
// pivots differ from partial pivoting; the factorization PA = LU is exact in
// the same sense, and its growth is close to partial pivoting in practice.
// The refinement of the mixed solver measures the result either way.
//
// The panel is W[r0 : r0+m, c0 : c0+w] (w <= 32).  Partial pivoting needs a
// reduction over all m rows for every column; here the rows are reduced in a
// tree of "playoffs" instead, and the workgroups meet once per leaf:
//   1. each wavefront holds up to 64*RPT0 rows in registers and picks w of
//      them by Gaussian elimination with partial pivoting on its own rows
//      (no interchanges: chosen rows just leave the competition);
//   2. wavefront 0 plays the 4*w winners of its workgroup the same way;
//   3. with more than one workgroup, every workgroup publishes its w
//      winners (row index and original values, as tagged words) and every
//      workgroup plays all G*w of them the same way (in two rounds when
//      G > 4), so all of them know the w pivot rows, in order, and the
//      factored w x w block [L11 \ U11] (the pivot rows as captured when
//      chosen in the last round);
//   4. each workgroup computes its own rows of L21 = A21 U11^-1 from their
//      original values (row-wise substitution with the operation order of
//      the elimination: multiply by 1/pivot, then the rank-1 updates, and
//      no scaling where |pivot| < eps), workgroup 0 writes [L11 \ U11] into
//      the top w rows and the rows those displaced into the winners' old
//      places, and turns the pivot order into libmat's sequential
//      interchanges (ipiv) and the leaf's row pairs.
// Every playoff plays on original row values, so no workgroup writes the
// matrix before all have read it: step 3 reads the published copies, and
// the rows a workgroup overwrites in step 4 are its own or the window's.
#include <algorithm>
#include <atomic>
#include <cstdint>

#include "ops.hpp"
#include "pivot.hpp"

namespace mg {

namespace {

constexpr int TS_NT = 256;  // threads per workgroup (4 wavefronts)
constexpr int TS_W = 32;    // widest leaf
constexpr int TS_RS = 36;   // LDS row stride in floats: 16-byte aligned rows
constexpr int TS_REC = 1 + TS_W;  // tagged words per published row: index, values

struct TsluP {
  float *W;
  long ld;
  int r0, m, c0, w;  // panel rows r0 .. r0+m-1 (r0 == c0), columns c0 .. c0+w-1
  float eps;
  int *ipiv;
  int *pairs;  // this leaf's row pairs (PAIR_STRIDE ints)
  int *err;    // set to 1 on a poll timeout
  uint64_t *rec;  // [G][TS_W][TS_REC] tagged words
  unsigned tag;   // distinct per launch, never 0
  int G;          // workgroups
};

// Orders the wavefront's LDS traffic around a lane-to-lane hand-off (one
// wavefront: the LDS serves its instructions in order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Wavefront max of non-negative int keys (-1: none): DPP within each
// 16-lane row (one v_max_i32 with a DPP operand per level), then the four
// row results on the scalar unit.
template <int CTRL>
__device__ __forceinline__ int dpp_maxi(int v) {
  // INT_MIN (the identity of max) as the old value lets the compiler fuse
  // the DPP move into the max; every lane of these patterns has a source
  return max(v, __builtin_amdgcn_update_dpp((int)0x80000000, v, CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ int wave_max_i32(int v) {
  v = dpp_maxi<0xB1>(v);   // quad_perm(1,0,3,2)
  v = dpp_maxi<0x4E>(v);   // quad_perm(2,3,0,1)
  v = dpp_maxi<0x141>(v);  // row_half_mirror
  v = dpp_maxi<0x140>(v);  // row_mirror
  const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return max(max(r0, r1), max(r2, r3));
}

// 1/x for the multipliers: the hardware reciprocal (1 ulp).  The playoffs
// and the substitution of the other rows use the same one, so the pivot
// rows and the other rows of L come from the same 1/pivot.
__device__ __forceinline__ float ts_recip(float x) { return __builtin_amdgcn_rcpf(x); }

// One chunk of four slots of a step's rank-1 update: the pivot row's values
// come from the winner lane (readlane) just before the update of the slot,
// so only four scalars are live at a time.
template <int RPT, int j, int k>
__device__ __forceinline__ void ts_chunk(float (&a)[RPT][TS_W], const float (&l)[RPT], int wq, int wl) {
  // slots j+1 .. of chunk k (none when the chunk ends at or before slot j)
  constexpr int s0 = 4 * k > j ? 4 * k : j + 1;
#pragma unroll
  for (int sl = s0; sl < 4 * k + 4; ++sl) {
    float src = a[0][sl];
#pragma unroll
    for (int q = 1; q < RPT; ++q) src = wq == q ? a[q][sl] : src;
    const float us = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(src), wl));
#pragma unroll
    for (int q = 0; q < RPT; ++q) a[q][sl] = fmaf(-us, l[q], a[q][sl]);
  }
}

// One elimination step of a playoff (see ts_play), column J = 4*g + j
// (slot j of the current group); slots < live_n hold unfinished columns.
template <int RPT, int j, bool KEEP>
__device__ __forceinline__ bool ts_step(float (&a)[RPT][TS_W], bool (&live)[RPT], int J, int live_n,
                                        float eps, int *win, float *rb, float *mult) {
  const int lane = threadIdx.x & 63;
  // keys: |value| bits (monotonic for non-negative floats), -1 for rows
  // out of the competition and for NaN
  int key[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const float v = a[q][j];
    key[q] = (live[q] && v == v) ? (__float_as_int(v) & 0x7fffffff) : -1;
  }
  int best = key[0];
#pragma unroll
  for (int q = 1; q < RPT; ++q) best = max(best, key[q]);
  const int mx = wave_max_i32(best);
  int wq = -1, wl = 0;
#pragma unroll
  for (int q = RPT - 1; q >= 0; --q) {
    const unsigned long long msk = mx >= 0 ? __ballot(key[q] == mx) : __ballot(live[q]);
    if (msk) {
      wq = q;
      wl = __builtin_ctzll(msk);
    }
  }
  if (wq < 0) return false;  // no rows left
  if (lane == 0) win[J] = wq * 64 + wl;
  float pvr = a[0][j];
#pragma unroll
  for (int q = 1; q < RPT; ++q) pvr = wq == q ? a[q][j] : pvr;
  const float pv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pvr), wl));
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (q == wq && lane == wl) {
      live[q] = false;
      if (KEEP) {
        // the row's current values as picked, in slot order (columns from J
        // on are exact; nothing in this playoff reads them back)
        float4 *row4 = reinterpret_cast<float4 *>(rb) + J * (TS_RS / 4);
#pragma unroll
        for (int i = 0; i < TS_W / 4; ++i)
          row4[i] = make_float4(a[q][4 * i], a[q][4 * i + 1], a[q][4 * i + 2], a[q][4 * i + 3]);
      }
    }
  }
  // |pivot| < eps: libmat's skip (no scaling, no update); the multiplier
  // then is the value itself
  const bool elim = !(fabsf(pv) < eps);
  const float inv = elim ? ts_recip(pv) : 1.0f;
  float l[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    l[q] = a[q][j] * inv;
    a[q][j] = l[q];
    if (KEEP) mult[(q * 64 + lane) * (TS_W + 1) + J] = l[q];
  }
  if (elim) {
    // chunks holding only finished columns are skipped (uniform branches;
    // independent ones, so every chunk updates its registers in place)
    if (4 * (j / 4) < live_n) ts_chunk<RPT, j, j / 4>(a, l, wq, wl);
    if (4 < live_n) ts_chunk<RPT, j, 1>(a, l, wq, wl);
    if (8 < live_n) ts_chunk<RPT, j, 2>(a, l, wq, wl);
    if (12 < live_n) ts_chunk<RPT, j, 3>(a, l, wq, wl);
    if (16 < live_n) ts_chunk<RPT, j, 4>(a, l, wq, wl);
    if (20 < live_n) ts_chunk<RPT, j, 5>(a, l, wq, wl);
    if (24 < live_n) ts_chunk<RPT, j, 6>(a, l, wq, wl);
    if (28 < live_n) ts_chunk<RPT, j, 7>(a, l, wq, wl);
  }
  return true;
}

// A playoff on the wavefront's candidate rows a[q][.] (RPT per lane;
// live[q]: a real row): Gaussian elimination with partial pivoting but no
// interchanges.  Step J picks the live row of largest |value| in column J
// (first (q, lane) on ties; NaN never wins, and a column of NaN takes the
// first live row), takes it out of the competition, and applies its
// multiplier and rank-1 update to every row (rows already out are updated
// too: only their values at the time they were picked are used).
// Columns go in groups of four: within a group every slot is static, and
// between groups the registers rotate by four, so group g holds column
// (4g + s) mod 32 in slot s while the loop over groups stays a runtime loop
// (fully unrolled steps would not fit the instruction cache).  Slots of
// finished columns take garbage updates; nothing reads them.
// Outputs: win[J] = the winner's slot q*64+lane (-1 once the rows run out)
// and, with KEEP, rb[J] = its row as picked in the slot order of its group
// (exact from column J on) and mult[slot][J] = every row's multiplier of
// column J (row stride TS_W + 1).
template <int RPT, bool KEEP>
__device__ __forceinline__ void ts_play(float (&a)[RPT][TS_W], bool (&live)[RPT], int w, float eps,
                                        int *win, float *rb, float *mult) {
  const int lane = threadIdx.x & 63;
  if (lane < TS_W) win[lane] = -1;
  wave_lds_sync();
#pragma unroll 1
  for (int g = 0; 4 * g < w; ++g) {
    const int J = 4 * g, live_n = w - 4 * g;  // slots < live_n: unfinished columns
    const bool go = ts_step<RPT, 0, KEEP>(a, live, J, live_n, eps, win, rb, mult) && J + 1 < w &&
                    ts_step<RPT, 1, KEEP>(a, live, J + 1, live_n, eps, win, rb, mult) && J + 2 < w &&
                    ts_step<RPT, 2, KEEP>(a, live, J + 2, live_n, eps, win, rb, mult) && J + 3 < w &&
                    ts_step<RPT, 3, KEEP>(a, live, J + 3, live_n, eps, win, rb, mult);
    (void)go;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const float t0 = a[q][0], t1 = a[q][1], t2 = a[q][2], t3 = a[q][3];
#pragma unroll
      for (int sl = 0; sl + 4 < TS_W; ++sl) a[q][sl] = a[q][sl + 4];
      a[q][TS_W - 4] = t0;
      a[q][TS_W - 3] = t1;
      a[q][TS_W - 2] = t2;
      a[q][TS_W - 1] = t3;
    }
  }
  wave_lds_sync();
}

// The factored pivot rows of a KEEP playoff in natural column order:
// Fn[J][c] = multiplier of column c < J (from mult, the winner's row), else
// the winner's value as picked (from rb).  One wavefront; w <= 32.
__device__ __forceinline__ void ts_factored_rows(const int *win, const float *rb, const float *mult,
                                                 int w, float *Fn) {
  const int lane = threadIdx.x & 63;
  for (int x = lane; x < w * TS_W; x += 64) {
    const int J = x / TS_W, c = x % TS_W;
    const int ws = win[J];
    float v = 0.0f;
    if (ws >= 0)
      v = c < J ? mult[ws * (TS_W + 1) + c] : rb[J * TS_RS + ((c - (J & ~3)) & (TS_W - 1))];
    Fn[J * TS_RS + c] = v;
  }
  wave_lds_sync();
}

// Polls one published row (TS_REC tagged words) until every word carries
// the tag; returns false on timeout.  Payloads go to out[0 .. TS_REC).
__device__ __forceinline__ bool ts_fetch(const uint64_t *src, unsigned tag, uint32_t *out) {
  uint64_t wd[TS_REC];
#pragma unroll
  for (int h = 0; h < TS_REC; ++h)
    wd[h] = __hip_atomic_load(src + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  long spins = 0;
  while (true) {
    bool ok = true;
#pragma unroll
    for (int h = 0; h < TS_REC; ++h) ok = ok && (uint32_t)(wd[h] >> 32) == tag;
    if (ok) break;
    if (++spins > SPIN_LIMIT) return false;
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int h = 0; h < TS_REC; ++h)
      if ((uint32_t)(wd[h] >> 32) != tag)
        wd[h] = __hip_atomic_load(src + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int h = 0; h < TS_REC; ++h) out[h] = (uint32_t)wd[h];
  return true;
}

// x := x U^-1 for NR rows held in registers, from the prepared pivot rows
// uc (LDS, row stride TS_RS): uc[J][0] = 1/u_JJ and uc[J][s] = u_J,J+s for
// the unfinished columns, zero elsewhere; a pivot with |u_JJ| < eps has
// uc[J][0] = 1 and a zero row, which reproduces the elimination's skip (no
// scaling, no update).  Operation order as in the elimination: x_J *= 1/u_JJ,
// then x_c -= x_J u_Jc.  Same register rotation as the playoff: on return
// column c < w is in slot (c - w) mod 32.
template <int NR>
__device__ __forceinline__ void ts_subst(float (&x)[NR][TS_W], const float *uc, int w) {
  const float4 *uc4 = reinterpret_cast<const float4 *>(uc);
#pragma unroll 1
  for (int J = 0; J < w; ++J) {
    float u[TS_W];
#pragma unroll
    for (int i = 0; i < TS_W / 4; ++i) {
      const float4 v = uc4[J * (TS_RS / 4) + i];
      u[4 * i] = v.x;
      u[4 * i + 1] = v.y;
      u[4 * i + 2] = v.z;
      u[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const float l = x[r][0] * u[0];
#pragma unroll
      for (int c = 1; c < TS_W; ++c) x[r][c - 1] = fmaf(-u[c], l, x[r][c]);
      x[r][TS_W - 1] = l;
    }
  }
}

__device__ __forceinline__ void ts_load_row(float (&x)[TS_W], const float *row) {
  const float4 *s = reinterpret_cast<const float4 *>(row);
#pragma unroll
  for (int i = 0; i < TS_W / 4; ++i) {
    const float4 v = s[i];
    x[4 * i] = v.x;
    x[4 * i + 1] = v.y;
    x[4 * i + 2] = v.z;
    x[4 * i + 3] = v.w;
  }
}

// RPT0: rows per lane in the first playoff (a workgroup holds 256*RPT0
// rows); RPTF: rows per lane in the first cross-workgroup round (G > 4).
template <int RPT0, int RPTF>
__global__ __launch_bounds__(TS_NT) void tslu_kernel(TsluP p) {
  constexpr int ROWS = TS_NT * RPT0;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, b = blockIdx.x;
  const int w = p.w, G = p.G;
  const int base = p.r0 + b * ROWS;
  const int rend = p.r0 + p.m;

  __shared__ __align__(16) float orig[ROWS * TS_RS];    // own rows, original values
  __shared__ __align__(16) float rb[4][TS_W * TS_RS];   // per-wavefront pivot rows
  __shared__ __align__(16) float cand[4 * TS_W * TS_RS];  // second-round candidates
  __shared__ float mult[2 * 64 * (TS_W + 1)];              // last playoff's multipliers
  __shared__ __align__(16) float Fn[TS_W * TS_RS];         // factored pivot rows
  __shared__ int win[4][TS_W];
  __shared__ int cid[4 * TS_W];  // candidate row indices of the current round
  __shared__ int fin[TS_W];      // pivot rows, in order
  __shared__ int s_pairs[PAIR_STRIDE];
  __shared__ unsigned char skip[ROWS];
  __shared__ int s_abort;

  if (t == 0) s_abort = 0;
  // 1. own rows: registers for the playoff, LDS for the original values
  float a[RPT0][TS_W];
  bool live[RPT0];
#pragma unroll
  for (int q = 0; q < RPT0; ++q) {
    const int li = (wave * RPT0 + q) * 64 + lane;
    const int row = base + li;
    live[q] = row < rend;
    const float *src = p.W + (long)row * p.ld + p.c0;
#pragma unroll
    for (int c = 0; c < TS_W; ++c) a[q][c] = (live[q] && c < w) ? src[c] : 0.0f;
    float4 *o = reinterpret_cast<float4 *>(orig + li * TS_RS);
#pragma unroll
    for (int i = 0; i < TS_W / 4; ++i)
      o[i] = make_float4(a[q][4 * i], a[q][4 * i + 1], a[q][4 * i + 2], a[q][4 * i + 3]);
    skip[li] = (row < p.r0 + w) ? 1 : 0;  // the window: written by workgroup 0
  }
  ts_play<RPT0, false>(a, live, w, p.eps, win[wave], rb[wave], mult);
  __syncthreads();

  // 2. the workgroup's playoff: wavefront 0 over the 4*w first-round winners
  if (wave == 0) {
    float a1[2][TS_W];
    bool l1[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ci = q * 64 + lane;
      const int v = ci / TS_W, j = ci % TS_W;
      const int s = j < w ? win[v][j] : -1;
      const int li = v * RPT0 * 64 + s;
      l1[q] = s >= 0;
      cid[ci] = s >= 0 ? base + li : -1;
      if (s >= 0) ts_load_row(a1[q], orig + li * TS_RS);
      else
#pragma unroll
        for (int c = 0; c < TS_W; ++c) a1[q][c] = 0.0f;
    }
    if (G == 1) {
      ts_play<2, true>(a1, l1, w, p.eps, win[0], rb[0], mult);
      ts_factored_rows(win[0], rb[0], mult, w, Fn);
    } else {
      ts_play<2, false>(a1, l1, w, p.eps, win[0], rb[0], mult);
    }
    if (lane < TS_W) {
      const int s = win[0][lane];
      fin[lane] = s >= 0 ? cid[s] : -1;
    }
  }
  __syncthreads();

  if (G > 1) {
    // 3a. publish the workgroup's winners: index and original values
    uint64_t *mine = p.rec + (long)b * TS_W * TS_REC;
    for (int x = t; x < TS_W * TS_REC; x += TS_NT) {
      const int j = x / TS_REC, h = x % TS_REC;
      const int row = fin[j];
      uint32_t pay;
      if (h == 0) pay = (uint32_t)row;
      else pay = row >= 0 ? __float_as_uint(orig[(row - base) * TS_RS + h - 1]) : 0u;
      __hip_atomic_store(mine + x, tagged(p.tag, pay), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();  // fin is rewritten below
    bool timeout = false;
    if (G <= 4) {
      // 3b. one round: wavefront 0 plays all G*w published rows
      if (wave == 0) {
        float a2[2][TS_W];
        bool l2[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int ci = q * 64 + lane;
          const int g = ci / TS_W, j = ci % TS_W;
          l2[q] = false;
          cid[ci] = -1;
#pragma unroll
          for (int c = 0; c < TS_W; ++c) a2[q][c] = 0.0f;
          if (g < G && j < w) {
            uint32_t wd[TS_REC];
            if (ts_fetch(p.rec + ((long)g * TS_W + j) * TS_REC, p.tag, wd)) {
              const int row = (int)wd[0];
              l2[q] = row >= 0;
              cid[ci] = row;
#pragma unroll
              for (int c = 0; c < TS_W; ++c) a2[q][c] = __uint_as_float(wd[1 + c]);
            } else {
              timeout = true;
            }
          }
        }
        if (__any(timeout)) {
          if (lane == 0) {
            s_abort = 1;
            __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        } else {
          ts_play<2, true>(a2, l2, w, p.eps, win[0], rb[0], mult);
          ts_factored_rows(win[0], rb[0], mult, w, Fn);
          if (lane < TS_W) {
            const int s = win[0][lane];
            fin[lane] = s >= 0 ? cid[s] : -1;
          }
        }
      }
    } else {
      // 3b. two rounds: wavefront v plays the rows of workgroups
      //     [v*K, v*K+K), then wavefront 0 plays the 4*w survivors
      const int K = (G + 3) / 4;
      const int g0 = wave * K;
      float a2[RPTF][TS_W];
      bool l2[RPTF];
#pragma unroll
      for (int q = 0; q < RPTF; ++q) {
        const int ci = q * 64 + lane;
        const int g = g0 + ci / TS_W, j = ci % TS_W;
        l2[q] = false;
#pragma unroll
        for (int c = 0; c < TS_W; ++c) a2[q][c] = 0.0f;
        if (ci < K * TS_W && g < G && j < w) {
          uint32_t wd[TS_REC];
          if (ts_fetch(p.rec + ((long)g * TS_W + j) * TS_REC, p.tag, wd)) {
            l2[q] = (int)wd[0] >= 0;
#pragma unroll
            for (int c = 0; c < TS_W; ++c) a2[q][c] = __uint_as_float(wd[1 + c]);
          } else {
            timeout = true;
          }
        }
      }
      if (__any(timeout)) {
        if (lane == 0) {
          s_abort = 1;
          __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      } else {
        ts_play<RPTF, false>(a2, l2, w, p.eps, win[wave], rb[wave], mult);
        // survivors' original values again from the published copies
        // (already complete), into the second-round candidates
        if (lane < TS_W) {
          const int s = lane < w ? win[wave][lane] : -1;
          const int ci = wave * TS_W + lane;
          float *dst = cand + ci * TS_RS;
          if (s >= 0) {
            const int g = g0 + s / TS_W, j = s % TS_W;
            const uint64_t *src = p.rec + ((long)g * TS_W + j) * TS_REC;
            cid[ci] = (int)(uint32_t)__hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int c = 0; c < TS_W; ++c)
              dst[c] = __uint_as_float((uint32_t)__hip_atomic_load(src + 1 + c, __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT));
          } else {
            cid[ci] = -1;
#pragma unroll
            for (int c = 0; c < TS_W; ++c) dst[c] = 0.0f;
          }
        }
      }
      __syncthreads();
      if (!s_abort && wave == 0) {
        float a3[2][TS_W];
        bool l3[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int ci = q * 64 + lane;
          l3[q] = cid[ci] >= 0;
          ts_load_row(a3[q], cand + ci * TS_RS);
        }
        ts_play<2, true>(a3, l3, w, p.eps, win[0], rb[0], mult);
        ts_factored_rows(win[0], rb[0], mult, w, Fn);
        if (lane < TS_W) {
          const int s = win[0][lane];
          fin[lane] = s >= 0 ? cid[s] : -1;
        }
      }
    }
    __syncthreads();
    if (s_abort) return;
  }

  // 4. every workgroup knows the pivot rows fin[0 .. w) and their factored
  //    values F (natural column order)
  const float *F = Fn;
  if (t < w) {
    const int row = fin[t];
    if (row >= base && row < base + ROWS) skip[row - base] = 1;  // moves into the window
  }
  if (b == 0 && wave == 0) {
    // libmat's sequential interchanges: step kk swaps row c0+kk with the
    // current position of pivot row kk
    int posw = lane < w ? fin[lane] : -1;
    int mypiv = 0;
    for (int kk = 0; kk < w; ++kk) {
      const int k = p.c0 + kk;
      const int pr = __builtin_amdgcn_readlane(posw, kk);
      if (lane == kk) mypiv = pr;
      if (posw == k) posw = pr;
      if (lane == kk) posw = k;
    }
    if (lane < w) p.ipiv[p.c0 + lane] = mypiv;
    swap_pairs(mypiv, p.c0, w, s_pairs);
  }
  // the pivot rows prepared for the substitution (into cand, free now)
  float *uc = cand;
  for (int x = t; x < TS_W * TS_W; x += TS_NT) {
    const int j = x / TS_W, sl = x % TS_W;
    if (j < w) {
      const float pv = F[j * TS_RS + j];
      const bool skipj = fabsf(pv) < p.eps;
      float v;
      if (sl == 0) v = skipj ? 1.0f : ts_recip(pv);
      else v = (!skipj && sl < w - j) ? F[j * TS_RS + j + sl] : 0.0f;
      uc[j * TS_RS + sl] = v;
    }
  }
  __syncthreads();
  // own rows: L21 = A21 U11^-1 from the original values
  {
    float x[RPT0][TS_W];
    bool mine[RPT0], any = false;
#pragma unroll
    for (int q = 0; q < RPT0; ++q) {
      const int li = (wave * RPT0 + q) * 64 + lane;
      ts_load_row(x[q], orig + li * TS_RS);
      mine[q] = base + li < rend && !skip[li];
      any = any || mine[q];
    }
    if (__any(any)) {
      ts_subst<RPT0>(x, uc, w);
#pragma unroll
      for (int q = 0; q < RPT0; ++q) {
        if (mine[q]) {
          float *dst = p.W + (long)(base + (wave * RPT0 + q) * 64 + lane) * p.ld + p.c0;
#pragma unroll
          for (int sl = 0; sl < TS_W; ++sl) {
            const int c = (sl + w) & (TS_W - 1);
            if (c < w) dst[c] = x[q][sl];
          }
        }
      }
    }
  }
  if (b == 0) {
    // the window gets the factored pivot rows
    for (int x = t; x < w * TS_W; x += TS_NT) {
      const int j = x / TS_W, c = x % TS_W;
      if (c < w) p.W[(long)(p.r0 + j) * p.ld + p.c0 + c] = F[j * TS_RS + c];
    }
    // window rows displaced into a pivot row's old place (outside the
    // window): their L21 rows from their original values
    const int np = s_pairs[0];
    if (wave == 0) {
      const int dst = lane < np ? s_pairs[1 + lane] : -1;
      const int src = lane < np ? s_pairs[1 + PAIR_MAX + lane] : -1;
      // (the source of such a move is always a window row; the bounds keep
      // the LDS read inside the workgroup's rows regardless)
      const bool out = dst >= p.r0 + w && src >= p.r0 && src < p.r0 + w;
      if (__any(out)) {
        float x[1][TS_W];
        ts_load_row(x[0], orig + (out ? src - p.r0 : 0) * TS_RS);
        ts_subst<1>(x, uc, w);
        if (out) {
          float *d = p.W + (long)dst * p.ld + p.c0;
#pragma unroll
          for (int sl = 0; sl < TS_W; ++sl) {
            const int c = (sl + w) & (TS_W - 1);
            if (c < w) d[c] = x[0][sl];
          }
        }
      }
    }
    // the leaf's row pairs for the later row moves
    for (int x = t; x < PAIR_STRIDE; x += TS_NT) p.pairs[x] = s_pairs[x];
  }
}

template <int RPT0, int RPTF>
void launch_tslu(hipStream_t s, const TsluP &p) {
  hipLaunchKernelGGL((tslu_kernel<RPT0, RPTF>), dim3(p.G), dim3(TS_NT), 0, s, p);
}

}  // namespace

// Largest panel height the tournament kernel takes (32 workgroups of 512 rows).
int tslu_max_rows() { return 32 * TS_NT * 2; }

// Tournament-pivoted factorization of the leaf W[c0:n, c0:c0+w] (w <= 32,
// n - c0 <= tslu_max_rows()): writes ipiv[c0 .. c0+w) (sequential
// interchanges, as libmat's) and the leaf's row pairs at
// pairs + c0 * PAIR_STRIDE.  The row moves outside the leaf's columns are
// the caller's (lu_rowswap).
void tslu_leaf(hipStream_t s, int n, float *W, long ld, int c0, int w, int *ipiv, int *pairs,
               float eps) {
  MG_CHECK(w >= 1 && w <= TS_W, "tslu_leaf: leaf width must be in [1, 32]");
  const int m = n - c0;
  MG_CHECK(m >= w && m <= tslu_max_rows(), "tslu_leaf: panel height out of range");
  // tags: distinct per launch, never 0; atomic so that host threads
  // factoring on different streams draw distinct values
  static std::atomic<unsigned> epoch{0};
  Workspace &ws = workspace(s);
  TsluP p{};
  p.W = W;
  p.ld = ld;
  p.r0 = c0;
  p.m = m;
  p.c0 = c0;
  p.w = w;
  p.eps = eps;
  p.ipiv = ipiv;
  p.pairs = pairs + (long)c0 * PAIR_STRIDE;
  p.err = sync_error_word(s);
  // a fresh allocation holds stale bytes, and a stale word whose high half
  // equals this launch's tag (small integers here) would pass as published:
  // clear it (tag 0 is never used) before the first kernel polls it
  const size_t rec_bytes = (size_t)32 * TS_W * TS_REC * 8;
  bool fresh = false;
  p.rec = (uint64_t *)ws.get("tslu.rec", rec_bytes, &fresh);
  if (fresh) fill_u32(s, (unsigned *)p.rec, (long)(rec_bytes / 4), 0u);
  unsigned tag = epoch.fetch_add(1, std::memory_order_relaxed) + 1;
  if (tag == 0) tag = epoch.fetch_add(1, std::memory_order_relaxed) + 1;
  p.tag = tag;
  const int rpt0 = m > 32 * TS_NT ? 2 : 1;
  p.G = ceil_div(m, TS_NT * rpt0);
  const int K = (p.G + 3) / 4;
  KScope ks(s, "panel", 2.0 * m * w * sizeof(float));
  if (rpt0 == 2) launch_tslu<2, 4>(s, p);
  else if (K <= 2) launch_tslu<1, 1>(s, p);
  else if (K <= 4) launch_tslu<1, 2>(s, p);
  else launch_tslu<1, 4>(s, p);
  MG_HIP(hipGetLastError());
}

}  // namespace mg
