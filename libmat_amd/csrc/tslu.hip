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
// tree of "playoffs" instead, and the workgroups meet once per tree level:
//   1. workgroup b holds rows r0 + 256b .. +255, one per lane of its four
//      wavefronts, and picks w of them by Gaussian elimination with partial
//      pivoting on its own rows (no interchanges: a chosen row leaves the
//      competition and is no longer updated, so it ends as its row of the
//      factored block);
//   2. the workgroups publish their winners' row indices and rows (the
//      values their playoff began with); the last of every group of eight
//      to arrive (an atomic ticket) loads the group's 256 winners (index
//      and row in one memory round trip) and plays them, and so on up the
//      tree (two
//      levels above the leaves for up to 16384 rows), until one workgroup
//      plays the final round, writes the factored w x w block [L11 \ U11]
//      and the pivot rows, and raises a flag;
//   3. every workgroup computes its rows of L21 = A21 U11^-1 from their
//      original values (row-wise substitution with the operation order of
//      the elimination: multiply by 1/pivot, then the rank-1 updates, and
//      no scaling where |pivot| < eps); workgroup 0 writes [L11 \ U11] into
//      the top w rows and the rows those displaced into the pivots' old
//      places and turns the pivot order into libmat's sequential
//      interchanges (ipiv) and the leaf's row pairs; two more workgroups,
//      launched past the ones holding rows, apply the interchanges to the
//      other columns of the block ([cl, ch)) and to perm, and write the
//      inverse of the unit lower 32 x 32 block for the block triangular
//      solves (trsm.hip).
// The second leaf of a 64-column node also does the node's update first:
// U12 = L11^-1 A12 from the first leaf's inverse, then every workgroup its
// rows' A22 -= L21 U12 in LDS before its leaf playoff (the tree nodes get
// the updated rows from the published winners, and every row of the leaf's
// columns is written once at the end).
// A playoff runs on all four SIMDs of a compute unit: every step, each
// wavefront finds its candidate (DPP max + ballot) and parks the
// candidate's row in LDS, one barrier, every wavefront takes the best of the
// four and the winner's row from LDS, and updates its own rows (about 1200
// cycles a step, measured; a playoff in one wavefront issues one instruction
// per four or more cycles, so 64 rows there cost about half as much per step
// as 256 rows here).  Rows per lane are a parameter (TS_RPT): four rows per
// lane took 2200 cycles a step, more than the level they save.
// Every playoff plays the rows' values as they were when the launch began
// (after the fused update): the leaf's columns are written only after the
// final round, so no workgroup overwrites a row another one may still read.
// An agent-scope release writes the L2 back, which the side stream's GEMM
// keeps dirty (a release fence in each of the four wavefronts instead of
// one cost ~7 us per leaf, measured), so a tree hand-off has one: the
// published rows are written by wavefront 0 alone, whose fence orders them
// before the ticket, which is itself relaxed.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <utility>

#include "ops.hpp"
#include "pivot.hpp"

namespace mg {

namespace {

constexpr int TS_NT = 256;   // threads per workgroup: 4 wavefronts
constexpr int TS_RPT = 1;    // playing rows per lane
constexpr int TS_PR = TS_NT * TS_RPT;  // rows (candidates) of one playoff
constexpr int TS_W = 32;     // widest leaf
constexpr int TS_RS = 36;    // LDS row stride in floats: 16-byte aligned rows
constexpr int TS_XS = 64;    // stride of the per-step exchange slots (a shift to address)
constexpr int TS_FAN = TS_PR / TS_W;  // tree nodes per node of the next level (8)
constexpr int TS_LMAX = 3;   // tree levels above the leaves: 256 -> 32 -> 4 -> 1
constexpr int TS_GMAX = 256;  // workgroups (one per 256 rows), all resident

struct TsluP {
  float *W;
  long ld;
  int r0, m, c0, w;  // panel rows r0 .. r0+m-1 (r0 == c0), columns c0 .. c0+w-1
  float eps;
  int *ipiv;
  int *pairs;     // this leaf's row pairs (PAIR_STRIDE ints)
  int *err;       // set to 1 on a poll timeout
  int *pub;       // [TS_LMAX][TS_GMAX][TS_W] winners' row indices per tree node
  unsigned *cnt;  // [TS_LMAX][TS_GMAX] arrival tickets (wrap to 0 after the last)
  float *prow;    // [TS_LMAX][TS_GMAX][TS_W][TS_W] winners' rows per tree node (the
                  // values their playoff began with; plain stores, ordered by the
                  // release fence before the ticket)
  float *F;       // [TS_W][TS_W] factored pivot rows of the final round
  int *fin;       // [TS_W] pivot rows (original indices), in order
  uint64_t *done; // tagged word: the final round is written
  unsigned tag;   // distinct per launch, never 0
  int G;          // workgroups holding rows (one per 256 rows); two more are launched
                  // for the tail (row moves, inverse)
  int P;          // leaf workgroups of the tree (one per TS_PR rows; = G here)
  int prio;       // wave priority (wave_prio of the launch stream)
  float *linv;    // if not null: the inverse of the leaf's unit lower 32 x 32
                  // block (identity past w), row-major at linv + c0 * 32
  int *perm;      // if not null: the row-index vector, moved with the rows
  int cl, ch;     // the leaf's interchanges also reach columns [cl, ch) outside it
  // second leaf of a 64-column node (prev_c0 >= 0): the node's first leaf
  // starts at prev_c0 (its inverse at linv + prev_c0 * 32); this launch also
  // does the node's U12 = L11^-1 A12 and A22 -= L21 U12 for its columns
  int prev_c0;
};

// Wavefront max of int keys, in every lane: DPP within each 16-lane row (one
// v_max_i32 with a DPP operand per level), then the rows combined by the
// gfx950 half-wave swaps.
template <int CTRL>
__device__ __forceinline__ int dpp_maxi(int v) {
  // INT_MIN (the identity of max) as the old value lets the compiler fuse
  // the DPP move into the max; every lane of these patterns has a source
  return max(v, __builtin_amdgcn_update_dpp((int)0x80000000, v, CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ int wave_max_all(int v) {
  v = dpp_maxi<0xB1>(v);   // quad_perm(1,0,3,2)
  v = dpp_maxi<0x4E>(v);   // quad_perm(2,3,0,1)
  v = dpp_maxi<0x141>(v);  // row_half_mirror
  v = dpp_maxi<0x140>(v);  // row_mirror
  const auto s32 = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  v = max((int)s32[0], (int)s32[1]);
  const auto s16 = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  return max((int)s16[0], (int)s16[1]);
}

// 1/x for the multipliers: the hardware reciprocal (1 ulp).  The playoffs
// and the substitution of the other rows use the same one, so the pivot
// rows and the other rows of L come from the same 1/pivot.
__device__ __forceinline__ float ts_recip(float x) { return __builtin_amdgcn_rcpf(x); }

typedef float ts_f4 __attribute__((ext_vector_type(4)));

// Words one workgroup hands to another within a launch (winners, the final
// round) go through agent-scope atomic loads and stores, as every exchange
// between workgroups in this library does: the compute units of different
// XCDs cache device memory in different L2s, and these accesses are the
// ones coherent across them without relying on cache maintenance.
__device__ __forceinline__ void ts_put(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ts_get(const int *p) {
  return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ts_putf(float *p, float v) {
  ts_put(reinterpret_cast<int *>(p), __float_as_int(v));
}
__device__ __forceinline__ float ts_getf(const float *p) {
  return __int_as_float(ts_get(reinterpret_cast<const int *>(p)));
}

// Per-step exchange of a workgroup playoff, double-buffered by step parity
// (one barrier per step: a buffer is rewritten two steps later, after every
// wavefront has passed the barrier that follows its reads).
struct __align__(16) TsXchg {
  float row[2][4][TS_XS];  // per wavefront: its candidate's row (columns J..)
  unsigned key[2][4];      // per wavefront: (rank << 10) | (1023 - slot)
};

// The playing rows of one lane: candidate slot q * 256 + thread (q < TS_RPT).
// a[q][] the row's current values (columns in natural order), live while
// the row may still win, pick = the step that chose it (-1: none).
struct TsRows {
  float a[TS_RPT][TS_W];
  bool live[TS_RPT];
  int pick[TS_RPT];
};

// The candidate's row (columns J..) into its wavefront's exchange slot.
template <int J, int Q>
__device__ __forceinline__ void ts_park(const TsRows &r, float *dst) {
#pragma unroll
  for (int c4 = J / 4; c4 < TS_W / 4; ++c4)
    *reinterpret_cast<float4 *>(dst + 4 * c4) =
        make_float4(r.a[Q][4 * c4], r.a[Q][4 * c4 + 1], r.a[Q][4 * c4 + 2], r.a[Q][4 * c4 + 3]);
}

template <int J, int... Q>
__device__ __forceinline__ void ts_parks(const TsRows &r, float *dst, bool cand, int wq,
                                         std::integer_sequence<int, Q...>) {
  ((cand && wq == Q ? ts_park<J, Q>(r, dst) : void()), ...);
}

// Step J of a workgroup playoff (FULL: w == 32, no step is skipped).  The
// candidate of a wavefront is its live row of largest |value| in column J
// (NaN never wins; a column of NaN takes the first live row), lowest slot
// on ties.  The workgroup's winner is the best of the four by a 22-bit
// rank, lowest slot on ties: the rank keeps the exponent and the top 14
// mantissa bits of |value|, so a candidate within a relative 2^-14 of
// another may win in its place (a pivot that large is as good for the
// elimination, and every wavefront takes the same decision).  The winner
// leaves the competition; every live row gets its multiplier (times
// 1/pivot) and the rank-1 update, both skipped where |pivot| < eps
// (libmat's rule).  Thread J's mywin gets the winner slot, -1 when no live
// rows are left.  qlive[q]: some row q of this wavefront was live when the
// playoff began (rows q of the wavefront that are all dead skip the update).
template <int J, bool FULL>
__device__ __forceinline__ void ts_step(TsRows &r, const bool (&qlive)[TS_RPT], int w, float eps,
                                        TsXchg &x, int &mywin) {
  if (!FULL && J >= w) return;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  constexpr int par = J & 1;
  int key[TS_RPT];
  int kbest = -1;
#pragma unroll
  for (int q = 0; q < TS_RPT; ++q) {
    const float v = r.a[q][J];
    key[q] = (r.live[q] && v == v) ? (__float_as_int(v) & 0x7fffffff) : -1;
    kbest = max(kbest, key[q]);
  }
  const int mx = wave_max_all(kbest);
  unsigned long long msk = 0;
  int wq = 0;
#pragma unroll
  for (int q = TS_RPT - 1; q >= 0; --q) {
    const unsigned long long m = __ballot(mx >= 0 ? key[q] == mx : r.live[q]);
    if (m) {
      msk = m;
      wq = q;
    }
  }
  const int wl = msk ? (int)__builtin_ctzll(msk) : 0;
  // the candidate parks its row: one guarded block per row index q (as
  // separate blocks, not alternatives of one branch, the four stay apart and
  // every register index stays static)
  ts_parks<J>(r, x.row[par][wave], msk != 0 && lane == wl, wq,
              std::make_integer_sequence<int, TS_RPT>{});
  if (lane == wl) {
    // the candidate (lane 0 when there is none) parks the key
    const unsigned rank = msk == 0 ? 0u : (mx >= 0 ? ((unsigned)mx >> 9) + 2u : 1u);
    x.key[par][wave] = (rank << 10) | (unsigned)(1023 - (wq * TS_NT + wave * 64 + wl));
  }
  __syncthreads();
  const uint4 kq = *reinterpret_cast<const uint4 *>(x.key[par]);
  // the best key, the same in every lane; the winner's LDS row address is
  // formed from it in vector registers (the scalar copy, which takes a
  // VALU-to-SALU round trip, only decides what is off the LDS chain)
  const unsigned bestv = max(max(kq.x, kq.y), max(kq.z, kq.w));
  const unsigned best = __builtin_amdgcn_readfirstlane(bestv);
  const bool found = (best >> 10) != 0;
  const int slot = 1023 - (int)(best & 1023u);
  const int wqq = slot >> 8, wlw = slot & 63;
  const int ww = (slot >> 6) & 3;
  const int wwv = (int)((~bestv >> 6) & 3u);  // = ww
  mywin = t == J ? (found ? slot : -1) : mywin;
  // the winner's row in whole 16-byte chunks; every load is issued first and
  // each chunk is waited for where its columns are used.  The first chunk
  // holds columns before J when J % 4 != 0: the empty asm marks them as
  // used, so it stays one ds_read_b128
  float u[TS_W];
  const ts_f4 *src =
      reinterpret_cast<const ts_f4 *>(&x.row[par][0][0] + wwv * TS_XS);
  ts_f4 qv[TS_W / 4];
#pragma unroll
  for (int c4 = J / 4; c4 < TS_W / 4; ++c4) qv[c4] = src[c4];
  if constexpr (J % 4 != 0) asm volatile("" : "+v"(qv[J / 4]));
#pragma unroll
  for (int c4 = J / 4; c4 < TS_W / 4; ++c4) {
    u[4 * c4] = qv[c4].x;
    u[4 * c4 + 1] = qv[c4].y;
    u[4 * c4 + 2] = qv[c4].z;
    u[4 * c4 + 3] = qv[c4].w;
  }
  const bool mine = found && wave == ww && lane == wlw;
  const float pv = u[J];
  const bool elim = found && !(fabsf(pv) < eps);
  const float inv = elim ? ts_recip(pv) : 1.0f;
#pragma unroll
  for (int q = 0; q < TS_RPT; ++q) {
    // a full playoff updates every row (lm = 0 where not live): no branch
    // between this update and the next step's reduction, so the two share
    // a basic block and their instructions can interleave
    if (!FULL && !qlive[q]) continue;
    const bool me = mine && q == wqq;
    r.pick[q] = me ? J : r.pick[q];
    r.live[q] = r.live[q] && !me;
    const float l = r.a[q][J] * inv;
    const float lm = (r.live[q] && elim) ? l : 0.0f;
    r.a[q][J] = r.live[q] ? l : r.a[q][J];
#pragma unroll
    for (int c = J + 1; c < TS_W; ++c) r.a[q][c] = fmaf(-u[c], lm, r.a[q][c]);
  }
}

template <bool FULL, int... J>
__device__ __forceinline__ void ts_steps(TsRows &r, const bool (&qlive)[TS_RPT], int w, float eps,
                                         TsXchg &x, int &mywin, std::integer_sequence<int, J...>) {
  (ts_step<J, FULL>(r, qlive, w, eps, x, mywin), ...);
}

// A playoff of the workgroup's TS_PR candidate slots (r.live marks the real
// ones); on return win[0 .. w) holds the winner slots.
__device__ __forceinline__ void ts_play(TsRows &r, int w, float eps, TsXchg &x, int *win) {
  bool qlive[TS_RPT];
#pragma unroll
  for (int q = 0; q < TS_RPT; ++q) {
    r.pick[q] = -1;
    qlive[q] = __ballot(r.live[q]) != 0;
  }
  int mywin = -1;
  if (w == TS_W)
    ts_steps<true>(r, qlive, w, eps, x, mywin, std::make_integer_sequence<int, TS_W>{});
  else
    ts_steps<false>(r, qlive, w, eps, x, mywin, std::make_integer_sequence<int, TS_W>{});
  if (threadIdx.x < TS_W) win[threadIdx.x] = mywin;
  __syncthreads();
}

// x := x U^-1 for the lane's row, from the final round's pivot rows in LDS,
// row J shifted to start at its diagonal: us[J][0] = 1/u_JJ (1 where
// |u_JJ| < eps) and us[J][k] = u_J,J+k for k >= 1 (zero where |u_JJ| < eps
// or J + k >= w), which reproduces the elimination: x_J *= 1/u_JJ, then
// x_c -= x_J u_Jc.  Rows J >= w are the identity's (x_J *= 1, no update),
// so every step runs and no branch keeps the loads of later steps from
// being issued early.
template <int J>
__device__ __forceinline__ void ts_sub(float (&x)[TS_W], const float *us) {
  constexpr int K = TS_W - J;  // entries J .. 31
  const float4 *row = reinterpret_cast<const float4 *>(us + J * TS_RS);
  float q[(K + 3) / 4 * 4];
#pragma unroll
  for (int k4 = 0; k4 < (K + 3) / 4; ++k4) {
    const float4 v = row[k4];
    q[4 * k4] = v.x;
    q[4 * k4 + 1] = v.y;
    q[4 * k4 + 2] = v.z;
    q[4 * k4 + 3] = v.w;
  }
  const float l = x[J] * q[0];
  x[J] = l;
#pragma unroll
  for (int k = 1; k < K; ++k) x[J + k] = fmaf(-q[k], l, x[J + k]);
}
template <int... J>
__device__ __forceinline__ void ts_subs(float (&x)[TS_W], const float *un,
                                        std::integer_sequence<int, J...>) {
  (ts_sub<J>(x, un), ...);
}

// Rows of the leaf between registers and memory: 16-byte vectors when every
// row start is aligned (vec), columns >= w as zeros.
__device__ __forceinline__ void ts_load_row(float (&x)[TS_W], const float *row, int w, bool vec) {
  if (vec && w == TS_W) {
#pragma unroll
    for (int c4 = 0; c4 < TS_W / 4; ++c4) {
      const float4 q = *reinterpret_cast<const float4 *>(row + 4 * c4);
      x[4 * c4] = q.x;
      x[4 * c4 + 1] = q.y;
      x[4 * c4 + 2] = q.z;
      x[4 * c4 + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < TS_W; ++c) x[c] = c < w ? row[c] : 0.0f;
  }
}
__device__ __forceinline__ void ts_store_row(float *row, const float (&x)[TS_W], int w, bool vec) {
  if (vec && w == TS_W) {
#pragma unroll
    for (int c4 = 0; c4 < TS_W / 4; ++c4)
      *reinterpret_cast<float4 *>(row + 4 * c4) =
          make_float4(x[4 * c4], x[4 * c4 + 1], x[4 * c4 + 2], x[4 * c4 + 3]);
  } else {
#pragma unroll
    for (int c = 0; c < TS_W; ++c)
      if (c < w) row[c] = x[c];
  }
}

__global__ __launch_bounds__(TS_NT) void tslu_kernel(TsluP p) {
  if (p.prio) __builtin_amdgcn_s_setprio(3);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, b = blockIdx.x;
  const int w = p.w;
  const int base = p.r0 + b * TS_NT;
  const int rend = p.r0 + p.m;
  const int myrow = base + t;
  const bool mine = myrow < rend;
  // 16-byte row accesses when the leaf's rows start 16-byte aligned
  const bool vec = ((p.ld & 3) == 0) && ((p.c0 & 3) == 0) &&
                   ((reinterpret_cast<uintptr_t>(p.W) & 15) == 0);

  __shared__ TsXchg xg;
  // own rows (original values), then the final pivot rows
  __shared__ __align__(16) float orig[(TS_NT + TS_W) * TS_RS];
  __shared__ __align__(16) float un[TS_W * TS_RS];  // final pivot rows, for the substitution
  __shared__ int cid[TS_PR];     // candidate row of every slot in the current playoff
  __shared__ int win[TS_W];      // winner slots of the current playoff
  __shared__ int fin[TS_W];      // pivot rows, in order
  __shared__ int s_fidx[TS_W];   // the final round's pivot rows, before publication
  __shared__ int windst[TS_W];   // where each window row goes (-1: into the window)
  __shared__ int s_pairs[PAIR_STRIDE];
  __shared__ unsigned mv_row[2][2 * TS_W];  // row moves: padded src, dst rows
  __shared__ unsigned char skip[TS_NT];
  __shared__ int s_go;  // this workgroup plays the next tree level
  __shared__ int s_ok;  // the final round arrived (no time-out)

  // 0. second leaf of a 64-column node: U12 = L11^-1 A12 of the node's
  //    first leaf (A12 = its window rows in this leaf's columns; the order of
  //    the block solve's X_0 = L_00^-1 B_0), kept in registers for the
  //    write-back and in LDS (the final pivot rows' place, unused until
  //    then) for the update of the own rows below
  const bool fused = p.prev_c0 >= 0;
  // the own row and, in a fused launch, its row of the first leaf's L21 are
  // loaded first, so that their memory round trip overlaps forming U12
  float x[TS_W], lp[TS_W];
  if (mine) ts_load_row(x, p.W + (long)myrow * p.ld + p.c0, w, vec);
  else
#pragma unroll
    for (int c = 0; c < TS_W; ++c) x[c] = 0.0f;
  if (fused && mine) ts_load_row(lp, p.W + (long)myrow * p.ld + p.prev_c0, TS_W, vec);
  float u4[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // U12 row t / 8, columns 4 (t % 8) ..
  float *U12s = orig + TS_NT * TS_RS;
  if (fused) {
    const float *lprev = p.linv + (long)p.prev_c0 * TS_W;
    {
      // all loads issued before any is used (one memory round trip)
      constexpr int FR = TS_W * TS_W / TS_NT;
      float lv[FR], av[FR];
#pragma unroll
      for (int j = 0; j < FR; ++j) {
        const int e = t + j * TS_NT, i = e / TS_W, c = e % TS_W;
        lv[j] = lprev[e];
        av[j] = c < w ? p.W[(long)(p.prev_c0 + i) * p.ld + p.c0 + c] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < FR; ++j) {
        const int e = t + j * TS_NT, i = e / TS_W, c = e % TS_W;
        un[i * TS_RS + c] = lv[j];
        U12s[i * TS_RS + c] = av[j];
      }
    }
    __syncthreads();
    const int i = t >> 3, cq = 4 * (t & 7);
#pragma unroll 8
    for (int j = 0; j < TS_W; ++j) {
      const float lv = un[i * TS_RS + j];
      const float4 a = *reinterpret_cast<const float4 *>(U12s + j * TS_RS + cq);
      u4[0] = fmaf(lv, a.x, u4[0]);
      u4[1] = fmaf(lv, a.y, u4[1]);
      u4[2] = fmaf(lv, a.z, u4[2]);
      u4[3] = fmaf(lv, a.w, u4[3]);
    }
    __syncthreads();
    *reinterpret_cast<float4 *>(U12s + i * TS_RS + cq) = make_float4(u4[0], u4[1], u4[2], u4[3]);
    __syncthreads();
  }

  // 1. own rows, original values, for the substitution (in a fused launch
  //    after A22 -= L21 U12)
  {
    if (fused && mine) {
      float acc[TS_W];
#pragma unroll
      for (int c = 0; c < TS_W; ++c) acc[c] = 0.0f;
#pragma unroll
      for (int j = 0; j < TS_W; ++j) {
#pragma unroll
        for (int c4 = 0; c4 < TS_W / 4; ++c4) {
          const float4 u = *reinterpret_cast<const float4 *>(U12s + j * TS_RS + 4 * c4);
          acc[4 * c4] = fmaf(lp[j], u.x, acc[4 * c4]);
          acc[4 * c4 + 1] = fmaf(lp[j], u.y, acc[4 * c4 + 1]);
          acc[4 * c4 + 2] = fmaf(lp[j], u.z, acc[4 * c4 + 2]);
          acc[4 * c4 + 3] = fmaf(lp[j], u.w, acc[4 * c4 + 3]);
        }
      }
#pragma unroll
      for (int c = 0; c < TS_W; ++c) x[c] = c < w ? x[c] - acc[c] : 0.0f;
    }
#pragma unroll
    for (int c4 = 0; c4 < TS_W / 4; ++c4)
      *reinterpret_cast<float4 *>(orig + t * TS_RS + 4 * c4) =
          make_float4(x[4 * c4], x[4 * c4 + 1], x[4 * c4 + 2], x[4 * c4 + 3]);
    skip[t] = 0;
  }
  // the own rows (original, or updated) are in LDS for the leaf playoff
  __syncthreads();

  // 2. the tree: the leaf workgroups play their rows, publish, and carry on as
  //    the next level's node when last to arrive (one playoff site, so its
  //    code is generated once)
  if (b < p.P) {
    TsRows r;
    static_assert(TS_RPT == 1, "the leaf playoff reads its rows from the LDS copy");
    {
      const int row = p.r0 + b * TS_PR + t;  // = myrow
      r.live[0] = row < rend;
      cid[t] = r.live[0] ? row : -1;
#pragma unroll
      for (int c4 = 0; c4 < TS_W / 4; ++c4) {
        const float4 q = *reinterpret_cast<const float4 *>(orig + t * TS_RS + 4 * c4);
        r.a[0][4 * c4] = q.x;
        r.a[0][4 * c4 + 1] = q.y;
        r.a[0][4 * c4 + 2] = q.z;
        r.a[0][4 * c4 + 3] = q.w;
      }
    }
    int level = 0, node = b, nodes = p.P;
    while (true) {
      const bool top = nodes == 1;
      ts_play(r, w, p.eps, xg, win);
      if (top) {
        // the final round: factored pivot rows and their indices, gathered
        // in LDS (the substitution's buffer, free until the tail) and
        // written by wavefront 0 behind one release (see above)
        static_assert(TS_RPT == 1, "one playing row per lane");
        if (r.pick[0] >= 0) {
#pragma unroll
          for (int c4 = 0; c4 < TS_W / 4; ++c4)
            *reinterpret_cast<float4 *>(un + r.pick[0] * TS_RS + 4 * c4) =
                make_float4(r.a[0][4 * c4], r.a[0][4 * c4 + 1], r.a[0][4 * c4 + 2],
                            r.a[0][4 * c4 + 3]);
          s_fidx[r.pick[0]] = cid[t];
        }
        __syncthreads();
        if (wave == 0) {
          if (lane < TS_W) {
            // every j < w has a winner (the final round has at least w rows)
            float4 *d = reinterpret_cast<float4 *>(p.F + lane * TS_W);
#pragma unroll
            for (int c4 = 0; c4 < TS_W / 4; ++c4)
              d[c4] = *reinterpret_cast<const float4 *>(un + lane * TS_RS + 4 * c4);
            p.fin[lane] = lane < w ? s_fidx[lane] : -1;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          if (lane == 0)
            __hip_atomic_store(p.done, tagged(p.tag, 1u), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      if (t < TS_W) {
        const int s = t < w ? win[t] : -1;
        ts_put(p.pub + ((long)level * TS_GMAX + node) * TS_W + t, s >= 0 ? cid[s] : -1);
      }
      if (t < w && win[t] >= 0) {
        // the winner's row for the next level, copied by wavefront 0 (whose
        // release fence below orders it): a leaf's candidates are its own
        // rows (LDS), a node's were read from the level below (still in
        // this XCD's L2)
        const int sw = win[t];
        float4 *d = reinterpret_cast<float4 *>(
            p.prow + (((long)level * TS_GMAX + node) * TS_W + t) * TS_W);
        float4 q[TS_W / 4];
        if (level == 0) {
#pragma unroll
          for (int c4 = 0; c4 < TS_W / 4; ++c4)
            q[c4] = *reinterpret_cast<const float4 *>(orig + sw * TS_RS + 4 * c4);
        } else {
          const float4 *src = reinterpret_cast<const float4 *>(
              p.prow + (((long)(level - 1) * TS_GMAX + node * TS_FAN + sw / TS_W) * TS_W +
                        sw % TS_W) * TS_W);
#pragma unroll
          for (int c4 = 0; c4 < TS_W / 4; ++c4) q[c4] = src[c4];
        }
#pragma unroll
        for (int c4 = 0; c4 < TS_W / 4; ++c4) d[c4] = q[c4];
      }
      const int parent = node / TS_FAN;
      const int members = min(TS_FAN, nodes - parent * TS_FAN);
      // wavefront 0 wrote everything the next level reads (indices and
      // rows): its release fence, then the ticket (relaxed: the fence
      // releases, the acquire fence below acquires; an acq_rel ticket
      // would write the L2 back a second time)
      if (wave == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) {
          const unsigned old = __builtin_amdgcn_atomic_inc32(
              p.cnt + (long)level * TS_GMAX + parent, (unsigned)(members - 1), __ATOMIC_RELAXED,
              "agent");
          s_go = old == (unsigned)(members - 1);
        }
      }
      __syncthreads();
      if (!s_go) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // this workgroup plays node (level + 1, parent): its members' winners,
      // member k's j-th winner in slot 32k + j; the index and the row are
      // loaded together (the row of an empty slot is not this launch's and
      // is dropped)
      const int m0 = parent * TS_FAN;
#pragma unroll
      for (int q = 0; q < TS_RPT; ++q) {
        const int sl = q * TS_NT + t;
        const int mem = sl / TS_W, j = sl % TS_W;
        const bool slot = mem < members && j < w;
        const long nd = (long)level * TS_GMAX + m0 + mem;
        int id = -1;
        if (slot) {
          id = ts_get(p.pub + nd * TS_W + j);
          ts_load_row(r.a[q], p.prow + (nd * TS_W + j) * TS_W, TS_W, true);
        }
        if (id < 0)
#pragma unroll
          for (int c = 0; c < TS_W; ++c) r.a[q][c] = 0.0f;
        cid[sl] = id;
        r.live[q] = id >= 0;
      }
      ++level;
      node = parent;
      nodes = (nodes + TS_FAN - 1) / TS_FAN;
    }
  }

  // 3. wait for the final round (a flag of its own: a wavefront that left
  //    the tree may not have read s_go yet when thread 0 gets here)
  if (t == 0) {
    long spins = 0;
    s_ok = 1;
    // relaxed polls (an acquire poll invalidates caches on every read); the
    // acquire fence after the barrier below orders what follows
    while (__hip_atomic_load(p.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
           tagged(p.tag, 1u)) {
      if (++spins > SPIN_LIMIT) {
        s_ok = 0;
        __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  if (!s_ok) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (fused && b == 0) {
    // U12 into the node's first leaf rows: every workgroup read A12 before
    // its leaf playoff, and the final round followed every leaf playoff
    const int i = t >> 3, cq = 4 * (t & 7);
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
      if (cq + cc < w) p.W[(long)(p.prev_c0 + i) * p.ld + p.c0 + cq + cc] = u4[cc];
  }
  // The tail is spread over the workgroups so that no one of them carries
  // all of it: workgroup 0 writes the window, ipiv and the row pairs (and
  // its rows of L21); two workgroups launched past the rows' G, which hold
  // no rows and sit out the tree, take the rest: G applies the interchanges
  // to the other columns and perm (~10 us at 16384 rows: the rows moved are
  // scattered over the panel), G + 1 forms the inverse of L11.  The launch
  // then ends with the longest of these instead of a sum (the row moves
  // followed by a workgroup's own L21 rows set its end before).
  const int mwg = p.G;
  const int lwg = p.G + 1;
  // the final pivot rows (factored) and their indices, once from memory
  // (every load issued before any is used: one memory round trip, where a
  // loop of load-then-store took one per element)
  float *Fs = orig + TS_NT * TS_RS;  // [TS_W][TS_RS]
  {
    static_assert(TS_W * TS_W % TS_NT == 0, "whole rounds of F");
    constexpr int FR = TS_W * TS_W / TS_NT;
    float fv[FR];
#pragma unroll
    for (int j = 0; j < FR; ++j) fv[j] = ts_getf(p.F + t + j * TS_NT);
    const int fi = t < w ? ts_get(p.fin + t) : -1;
#pragma unroll
    for (int j = 0; j < FR; ++j) {
      const int e = t + j * TS_NT;
      Fs[(e / TS_W) * TS_RS + e % TS_W] = fv[j];
    }
    if (t < TS_W) {
      fin[t] = fi;
      windst[t] = -1;
    }
  }
  __syncthreads();
  if ((b == 0 || b == mwg) && wave == 0) {
    // libmat's sequential interchanges: step kk swaps row c0+kk with the
    // current position of pivot row kk.  Lane i < w follows window position
    // c0+i, lane 32+j the original place of pivot row j when that lies below
    // the window (the only places the swaps touch): occ = the original row
    // there now.
    const int fj = lane >= TS_W && lane - TS_W < w ? fin[lane - TS_W] : -1;
    int pos = -1;
    if (lane < w) pos = p.c0 + lane;
    else if (fj >= p.c0 + w) pos = fj;
    int occ = pos;
    const int finl = lane < w ? fin[lane] : -1;
    int mypiv = 0;
    for (int kk = 0; kk < w; ++kk) {
      const int f = __builtin_amdgcn_readlane(finl, kk);
      const unsigned long long m = __ballot(pos >= 0 && occ == f);
      const int sl = (int)__builtin_ctzll(m);
      const int okk = __builtin_amdgcn_readlane(occ, kk);
      const int osl = __builtin_amdgcn_readlane(occ, sl);
      const int psl = __builtin_amdgcn_readlane(pos, sl);
      mypiv = lane == kk ? psl : mypiv;
      occ = lane == kk ? osl : (lane == sl ? okk : occ);
    }
    if (b == 0 && lane < w) p.ipiv[p.c0 + lane] = mypiv;
    // the row pairs (dst receives original row src) and where each window
    // row goes: a moved window row below the window is displaced, one inside
    // it is a pivot row (no L21 row)
    const bool moved = pos >= 0 && occ != pos;
    const unsigned long long mm = __ballot(moved);
    if (moved) {
      const int at = __popcll(mm & ((1ull << lane) - 1ull));
      s_pairs[1 + at] = pos;
      s_pairs[1 + PAIR_MAX + at] = occ;
      if (pos >= p.c0 + w && occ < p.c0 + w) windst[occ - p.c0] = pos;
    }
    if (lane == 0) s_pairs[0] = __popcll(mm);
  }
  // prepared for the substitution (see ts_sub; entries past a row's end are
  // read but never used)
  for (int e = t; e < TS_W * TS_W; e += TS_NT) {
    const int J = e / TS_W, c = e % TS_W;
    if (c >= J) {
      const float pjj = Fs[J * TS_RS + J];
      const bool skipj = !(J < w) || fabsf(pjj) < p.eps;
      float v = 0.0f;
      if (c == J) v = skipj ? 1.0f : ts_recip(pjj);
      else if (c < w && !skipj) v = Fs[J * TS_RS + c];
      un[J * TS_RS + (c - J)] = v;
    }
  }
  if (t < w) {
    const int row = fin[t];
    if (row >= base && row < base + TS_NT) skip[row - base] = 1;  // moves into the window
  }
  __syncthreads();

  if (b == mwg) {
    // the interchanges reach the other columns [cl, ch) (one thread per
    // column: all of its reads, then all of its writes) and perm (read
    // before, written after the columns).  Pairs past np copy row src[0]
    // onto itself and are stored first, so the loops run without branches:
    // every load is issued before any store.  The columns outside the leaf
    // are not touched by anything else in this launch.
    // (The padded rows sit in LDS and are read where they are used: held
    // in registers they lifted the kernel past 256 VGPRs, which kept it off
    // compute units running two split-GEMM workgroups.)
    const int np = s_pairs[0];
    const int pv = (p.perm && t < np) ? p.perm[s_pairs[1 + PAIR_MAX + t]] : 0;
    const bool inside = p.c0 >= p.cl && p.c0 + w <= p.ch;  // the leaf's own columns are skipped
    const int nother = (p.ch - p.cl) - (inside ? w : 0);
    if (t < 2 * TS_W) {
      const int s0 = s_pairs[1 + PAIR_MAX];
      mv_row[0][t] = (unsigned)(t < np ? s_pairs[1 + PAIR_MAX + t] : s0);
      mv_row[1][t] = (unsigned)(t < np ? s_pairs[1 + t] : s0);
    }
    __syncthreads();
    const unsigned long uld = (unsigned long)p.ld;
    for (int o = t; o < nother; o += TS_NT) {
      const int c = p.cl + o + ((inside && p.cl + o >= p.c0) ? w : 0);
      float *Wc = p.W + c;
      float v[2 * TS_W];
#pragma unroll
      for (int x = 0; x < 2 * TS_W; ++x) v[x] = Wc[mv_row[0][x] * uld];
#pragma unroll
      for (int x = 2 * TS_W - 1; x >= 0; --x) Wc[mv_row[1][x] * uld] = v[x];
    }
    if (p.perm) {
      __syncthreads();
      if (t < np) p.perm[s_pairs[1 + t]] = pv;
    }
  }
  if (b == lwg && wave == 1 && p.linv && lane < TS_W) {
    // column `lane` of the inverse of the unit lower factored block by
    // forward substitution (rows past w: identity), for the block triangular
    // solves with this leaf's L11 (trsm.hip).  Column-oriented: once x_i is
    // final it updates the rows below it, so row j still takes its terms in
    // the order k = 0 .. j-1, and the steps are independent FMAs
    float xi[TS_W];
#pragma unroll
    for (int i = 0; i < TS_W; ++i) xi[i] = i == lane ? 1.0f : 0.0f;
#pragma unroll
    for (int i = 0; i < TS_W - 1; ++i) {
#pragma unroll
      for (int j = i + 1; j < TS_W; ++j) {
        const float f = j < w ? Fs[j * TS_RS + i] : 0.0f;
        xi[j] = j < w ? fmaf(-f, xi[i], xi[j]) : xi[j];
      }
    }
    float *dst = p.linv + (long)p.c0 * TS_W + lane;
#pragma unroll
    for (int i = 0; i < TS_W; ++i) dst[i * TS_W] = xi[i];
  }

  // 4. L21 = A21 U11^-1 from the original values, for the own rows below
  //    the window that stay, and in workgroup 0 for the window rows, which
  //    either move into the window as pivot rows or to a pivot row's old
  //    place
  const bool winrow = b == 0 && t < w;
  const bool stays = mine && !skip[t] && myrow >= p.r0 + w;
  if (stays || winrow) {
    float x[TS_W];
#pragma unroll
    for (int c4 = 0; c4 < TS_W / 4; ++c4) {
      const float4 q = *reinterpret_cast<const float4 *>(orig + t * TS_RS + 4 * c4);
      x[4 * c4] = q.x;
      x[4 * c4 + 1] = q.y;
      x[4 * c4 + 2] = q.z;
      x[4 * c4 + 3] = q.w;
    }
    ts_subs(x, un, std::make_integer_sequence<int, TS_W>{});
    if (stays) ts_store_row(p.W + (long)myrow * p.ld + p.c0, x, w, vec);
    if (winrow && windst[t] >= 0) ts_store_row(p.W + (long)windst[t] * p.ld + p.c0, x, w, vec);
  }
  if (b == 0) {
    // the window gets the factored pivot rows
    for (int e = t; e < w * TS_W; e += TS_NT) {
      const int J = e / TS_W, c = e % TS_W;
      if (c < w) p.W[(long)(p.r0 + J) * p.ld + p.c0 + c] = Fs[J * TS_RS + c];
    }
    // the leaf's row pairs for the later row moves
    for (int e = t; e < PAIR_STRIDE; e += TS_NT) p.pairs[e] = s_pairs[e];
  }
}

}  // namespace

// Largest panel height the tournament kernel takes (one workgroup per 256
// rows, all of them resident at once).
int tslu_max_rows() { return TS_GMAX * TS_NT; }

// Tournament-pivoted factorization of the leaf W[c0:n, c0:c0+w] (w <= 32,
// n - c0 <= tslu_max_rows()): writes ipiv[c0 .. c0+w) (sequential
// interchanges, as libmat's) and the leaf's row pairs at
// pairs + c0 * PAIR_STRIDE, and applies the interchanges to the columns
// [cl, ch) outside the leaf and to perm (if not null).  linv (if not null)
// receives the inverse of the leaf's unit lower block at linv + c0 * 32;
// prev_c0 >= 0 makes it the second leaf of a 64-column node (see the
// kernel's comment), which needs linv.
void tslu_leaf(hipStream_t s, int n, float *W, long ld, int c0, int w, int *ipiv, int *pairs,
               float eps, float *linv, int *perm, int cl, int ch, int prev_c0) {
  MG_CHECK(prev_c0 < 0 || (linv && prev_c0 + TS_W == c0),
           "tslu_leaf: a fused leaf follows its node's first leaf, with the inverses");
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
  // one slot: tickets (must start at zero; every complete launch leaves them
  // at zero again), the done word (a stale word whose high half equals this
  // launch's tag would pass as raised: cleared when fresh, tag 0 is never
  // used), published winners and the final round
  const size_t cnt_bytes = (size_t)TS_LMAX * TS_GMAX * sizeof(unsigned);
  const size_t done_off = cnt_bytes;
  const size_t pub_off = done_off + 256;
  const size_t pub_bytes = (size_t)TS_LMAX * TS_GMAX * TS_W * sizeof(int);
  const size_t f_off = pub_off + pub_bytes;
  const size_t fin_off = f_off + (size_t)TS_W * TS_W * sizeof(float);
  const size_t prow_off = fin_off + 256;
  const size_t bytes = prow_off + (size_t)TS_LMAX * TS_GMAX * TS_W * TS_W * sizeof(float);
  bool fresh = false;
  char *buf = (char *)ws.get("tslu.ws", bytes, &fresh);
  if (fresh) fill_u32(s, (unsigned *)buf, (long)(pub_off / 4), 0u);
  p.cnt = (unsigned *)buf;
  p.done = (uint64_t *)(buf + done_off);
  p.pub = (int *)(buf + pub_off);
  p.F = (float *)(buf + f_off);
  p.fin = (int *)(buf + fin_off);
  p.prow = (float *)(buf + prow_off);
  unsigned tag = epoch.fetch_add(1, std::memory_order_relaxed) + 1;
  if (tag == 0) tag = epoch.fetch_add(1, std::memory_order_relaxed) + 1;
  p.tag = tag;
  p.G = ceil_div(m, TS_NT);
  p.P = ceil_div(m, TS_PR);
  p.prio = wave_prio(s);
  p.linv = linv;
  p.perm = perm;
  p.cl = cl;
  p.ch = ch;
  p.prev_c0 = prev_c0;
  KScope ks(s, "panel", 2.0 * m * w * sizeof(float));
  hipLaunchKernelGGL(tslu_kernel, dim3(p.G + 2), dim3(TS_NT), 0, s, p);
  MG_HIP(hipGetLastError());
}

}  // namespace mg
