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
// MI355X design.  The panel's rows are spread over P <= 64 workgroups of
// 256 threads, each thread holding one (or, for m > 16384, two) whole rows of
// the leaf in registers for the whole leaf (register rotation keeps the
// current column in slot 0, so every register index is static while the
// column loop stays a runtime loop).  One column step:
//   * each wavefront finds its best row (f32: a DPP float max, then ballots
//     for the first row holding it; f64: 64-bit keys) and parks that row,
//     plus row k if it owns it, in LDS (double-buffered by step parity);
//   * one barrier; wavefront 0 merges the wavefront candidates lane-parallel;
//   * with P > 1, wavefront 0 publishes the workgroup's candidate (value,
//     row index, row) to global memory as self-validating 64-bit words
//     (payload in the low half, a tag unique to this launch and step in the
//     high half), polls the other workgroups' records until every tag
//     matches, reduces the P candidates (same data, same answer in every
//     workgroup) and parks the pivot row and row k in LDS; one more barrier.
//     No fences, atomics or counters are needed and nothing is reset between
//     launches: stale words carry an older tag.
//   * swap; scale and rank-1 update of the unfinished columns, fused with
//     the rotation and free of branches.
// At the end, workgroup 0 turns the leaf's interchanges into row pairs that
// every later row move of this leaf reuses.  Every poll is bounded and
// reports a timeout instead of hanging.
//
// Two variants of the exchange (panel_exchange() picks per type):
//   * LOCAL: every workgroup of the launch on one XCD (8x the workgroups,
//     every eighth takes part, dispatch being round-robin over the XCDs);
//     the words are stored at workgroup scope and stay in that XCD's L2,
//     where the other workgroups' agent-scope polls find them;
//   * panel_gepp_la_kernel: the rank-1 update of a step runs while the next
//     step's keys travel; only the next column is updated first.
#include <algorithm>
#include <atomic>
#include <cstdint>

#include "ops.hpp"
#include "pivot.hpp"

namespace mg {

// threads per panel workgroup.  Fewer wavefronts per barrier beat fewer
// workgroups per exchange; measured per column at m = 16384 (f32): 1024
// threads 4.64 us, 512 threads 4.03 us, 256 threads 3.72 us.
template <typename T>
struct PanelNT {
  static constexpr int value = 256;
};

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
  unsigned epoch;        // distinct per launch (tags)
  uint64_t *rec;         // [2][nwg][REC] tagged words: value, row, row values
  uint64_t *krec;        // [2][KREC] tagged words: row k
  int *err;              // set to 1 on a poll timeout
  int *pairs;            // this leaf's interchanges as row pairs (see swap_pairs)
  int prio;              // wave priority (wave_prio of the launch stream)
};
template <typename T>
__device__ inline void better(T &bv, int &bi, T v, int i) {
  if (v > bv || (v == bv && i < bi)) {
    bv = v;
    bi = i;
  }
}

// A value of type T as 32-bit payload words.
template <typename T>
struct Words {
  static constexpr int N = sizeof(T) / 4;
};

// Wavefront max of 64-bit keys: DPP within each 16-lane row, readlane
// across the four rows; every lane gets the result.
template <int CTRL>
__device__ inline uint64_t dpp_max(uint64_t k) {
  const uint32_t lo = __builtin_amdgcn_update_dpp((uint32_t)k, (uint32_t)k, CTRL, 0xf, 0xf, false);
  const uint32_t hi =
      __builtin_amdgcn_update_dpp((uint32_t)(k >> 32), (uint32_t)(k >> 32), CTRL, 0xf, 0xf, false);
  const uint64_t o = ((uint64_t)hi << 32) | lo;
  return o > k ? o : k;
}

__device__ inline uint64_t wave_max_u64(uint64_t k) {
  k = dpp_max<0xB1>(k);   // quad_perm(1,0,3,2)
  k = dpp_max<0x4E>(k);   // quad_perm(2,3,0,1)
  k = dpp_max<0x141>(k);  // row_half_mirror
  k = dpp_max<0x140>(k);  // row_mirror
  uint64_t r = 0;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)k, row * 16);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(k >> 32), row * 16);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    r = o > r ? o : r;
  }
  return r;
}

// Pivot candidates as keys: larger key = larger |value|, then lower row.
// Key 0 means "no candidate" (row not eligible, or NaN, which the
// reference's strict '>' scan never selects).  f32: (|v| bits, ~row) in one
// 64-bit key.  f64: the value key (|v| bits + 1) and the row key (~row) are
// reduced one after the other.
template <typename T>
struct PivKey;
template <>
struct PivKey<float> {
  uint64_t k;
  __device__ void clear() { k = 0; }
  __device__ void offer(float v, int row) {
    const float a = fabsf(v);
    if (!(a == a)) return;
    const uint64_t c = ((uint64_t)__float_as_uint(a) << 32) | (uint32_t)~(uint32_t)row;
    if (c > k) k = c;
  }
  __device__ void wave_reduce() { k = wave_max_u64(k); }
  __device__ void merge(const PivKey &o) {
    if (o.k > k) k = o.k;
  }
  __device__ bool valid() const { return k != 0; }
  __device__ int row() const { return (int)~(uint32_t)k; }
  static constexpr int NW = 2;  // 32-bit payload words
  __device__ uint32_t word(int i) const { return i == 0 ? (uint32_t)k : (uint32_t)(k >> 32); }
  __device__ void set_word(int i, uint32_t w) {
    k = i == 0 ? ((k & 0xffffffff00000000ull) | w) : ((k & 0xffffffffull) | ((uint64_t)w << 32));
  }
};
template <>
struct PivKey<double> {
  uint64_t v;  // |value| bits + 1
  uint32_t r;  // ~row
  __device__ void clear() {
    v = 0;
    r = 0;
  }
  __device__ void offer(double x, int row) {
    const double a = fabs(x);
    if (!(a == a)) return;
    const uint64_t c = (uint64_t)__double_as_longlong(a) + 1ull;
    const uint32_t cr = ~(uint32_t)row;
    if (c > v || (c == v && cr > r)) {
      v = c;
      r = cr;
    }
  }
  __device__ void wave_reduce() {
    const uint64_t vm = wave_max_u64(v);
    r = (uint32_t)wave_max_u64(v == vm ? (uint64_t)r : 0ull);
    v = vm;
  }
  __device__ void merge(const PivKey &o) {
    if (o.v > v || (o.v == v && o.r > r)) {
      v = o.v;
      r = o.r;
    }
  }
  __device__ bool valid() const { return v != 0; }
  __device__ int row() const { return (int)~r; }
  static constexpr int NW = 3;
  __device__ uint32_t word(int i) const {
    return i == 0 ? (uint32_t)v : (i == 1 ? (uint32_t)(v >> 32) : r);
  }
  __device__ void set_word(int i, uint32_t w) {
    if (i == 0) v = (v & 0xffffffff00000000ull) | w;
    else if (i == 1) v = (v & 0xffffffffull) | ((uint64_t)w << 32);
    else r = w;
  }
};


template <typename T>
__device__ inline uint32_t word_of(T v, int h) {
  if constexpr (sizeof(T) == 4) {
    return __float_as_uint(v);
  } else {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return h ? (uint32_t)(u >> 32) : (uint32_t)u;
  }
}

template <typename T>
__device__ inline T value_of(const uint32_t *w) {
  if constexpr (sizeof(T) == 4) {
    return __uint_as_float(w[0]);
  } else {
    const unsigned long long u = ((unsigned long long)w[1] << 32) | w[0];
    return __longlong_as_double((long long)u);
  }
}

// A row of N values between registers and a 16-byte aligned LDS row, in
// 16-byte accesses.
template <typename T, int N>
__device__ __forceinline__ void lds_store_row(T *dst, const T (&v)[N]) {
  if constexpr (sizeof(T) == 4) {
    float4 *d = reinterpret_cast<float4 *>(dst);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) d[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
  } else {
    double2 *d = reinterpret_cast<double2 *>(dst);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) d[i] = make_double2(v[2 * i], v[2 * i + 1]);
  }
}
template <typename T, int N>
__device__ __forceinline__ void lds_load_row(const T *src, T (&v)[N]) {
  if constexpr (sizeof(T) == 4) {
    const float4 *s4 = reinterpret_cast<const float4 *>(src);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
      const float4 x = s4[i];
      v[4 * i] = x.x, v[4 * i + 1] = x.y, v[4 * i + 2] = x.z, v[4 * i + 3] = x.w;
    }
  } else {
    const double2 *s2 = reinterpret_cast<const double2 *>(src);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const double2 x = s2[i];
      v[2 * i] = x.x, v[2 * i + 1] = x.y;
    }
  }
}

// Exchange stores.  Across the whole device the words are written at agent
// scope (through to memory, where every XCD's loads find them).  With all
// of a launch's workgroups on one XCD (LOCAL), they are written at
// workgroup scope and stay in that XCD's L2, where the agent-scope loads of
// the other workgroups on the same XCD find them (measured on MI355X: a
// store/poll round trip of ~0.57 us instead of ~1.07 us).
template <bool LOCAL>
__device__ __forceinline__ void xstore(uint64_t *p, uint64_t v) {
  if constexpr (LOCAL) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i on
// XCD i mod 8).  A LOCAL launch has 8x the workgroups and only every
// eighth takes part, so all of them share XCD 0 and its L2.
constexpr int NXCD = 8;
template <bool LOCAL>
__device__ __forceinline__ int panel_wg() {
  return LOCAL ? (int)(blockIdx.x / NXCD) : (int)blockIdx.x;
}
template <bool LOCAL>
__device__ __forceinline__ bool panel_idle() {
  return LOCAL && (blockIdx.x % NXCD) != 0;
}

template <typename T, int WMAX, int RPT, bool LOCAL>
__global__ __launch_bounds__(PanelNT<T>::value) void panel_gepp_kernel(PanelP<T> p) {
  if (p.prio) __builtin_amdgcn_s_setprio(3);
  constexpr int PANEL_NT = PanelNT<T>::value;
  constexpr int PANEL_NW = PANEL_NT / 64;
  constexpr int RPW = PANEL_NT * RPT;
  constexpr int NWD = Words<T>::N;
  // record words: [0..KW) key, [ROW0..) row
  constexpr int KW = PivKey<T>::NW;
  constexpr int ROW0 = 4;
  constexpr int REC = ROW0 + WMAX * NWD;
  constexpr int KREC = WMAX * NWD;
  if (panel_idle<LOCAL>()) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int b = panel_wg<LOCAL>();
  const int P = p.nwg;
  const int rend = p.r0 + p.m;
  const int w = p.w;

  __shared__ __align__(16) T wrow[2][PANEL_NW][WMAX];  // each wavefront's candidate row
  __shared__ __align__(16) PivKey<T> wkey[2][PANEL_NW];
  __shared__ __align__(16) T krow[2][WMAX];            // row k (if this workgroup owns it)
  __shared__ __align__(16) uint32_t stage[2][2 * WMAX * NWD];  // winner row, row k (payload words)
  __shared__ int s_piv[2], s_gw[2];
  __shared__ int s_hist[WMAX];  // pivot row of every step (k itself when not pivoted)
  __shared__ int s_abort;

  // Register rotation: at step j, slot 0 holds column j, slots 1..w-1-j the
  // columns still to factor, and finished columns rotate to the tail (whole
  // rows move on an interchange).  After w steps column c is in slot
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
      const int par = step & 1;
      // 1. wavefront candidate (first row of max |a|); park it and row k
      PivKey<T> key;
      key.clear();
      int wbest = -1;
      if constexpr (sizeof(T) == 4) {
        // float max over the wavefront, then the first (slot, lane) holding it:
        // rows increase with the slot q, then with the lane
        float tv = -1.0f;
        int tq = -1;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const float av = fabsf(a[q][0]);
          if (row[q] < rend && row[q] >= k && av > tv) {  // NaN never wins
            tv = av;
            tq = q;
          }
        }
        const float mx = wave_max_f32(tv);
        if (mx >= 0.0f) {
#pragma unroll
          for (int q = RPT - 1; q >= 0; --q) {
            const unsigned long long msk = __ballot(tq == q && tv == mx);
            if (msk) wbest = p.r0 + b * RPW + q * PANEL_NT + wave * 64 + __builtin_ctzll(msk);
          }
          key.offer(mx, wbest);
        }
      } else {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (row[q] < rend && row[q] >= k) key.offer(a[q][0], row[q]);
        key.wave_reduce();
        wbest = key.valid() ? key.row() : -1;
      }
      if (lane == 0) wkey[par][wave] = key;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (row[q] == wbest) lds_store_row<T, WMAX>(wrow[par][wave], a[q]);
        if (row[q] == k) lds_store_row<T, WMAX>(krow[par], a[q]);
      }
      __syncthreads();
      // 2. workgroup candidate, lane-parallel in wavefront 0: lane x < 8
      //    holds wavefront x's key; the pivot index (and for P == 1 the
      //    wavefront that parked the pivot row) goes to LDS
      int piv;
      const T *win_row = nullptr;
      const T *k_row = nullptr;
      if (P == 1) {
        if (wave == 0) {
          PivKey<T> mk;
          mk.clear();
          if (lane < PANEL_NW) mk = wkey[par][lane];
          PivKey<T> gkey = mk;
          gkey.wave_reduce();
          const unsigned long long gm =
              __ballot(lane < PANEL_NW && mk.valid() && gkey.valid() && mk.row() == gkey.row());
          if (lane == 0) {
            s_piv[par] = gkey.valid() ? gkey.row() : k;
            s_gw[par] = gm ? __builtin_ctzll(gm) : 0;
          }
        }
        __syncthreads();
        piv = s_piv[par];
      } else {
        // 3. exchange between workgroups (wavefront 0 of every workgroup):
        //    publish this workgroup's key and row (and row k by its owner),
        //    poll the P keys, then fetch the winner's row and row k
        const unsigned tag = p.epoch * 64u + (unsigned)step + 1u;
        if (wave == 0) {
          PivKey<T> mk;
          mk.clear();
          if (lane < PANEL_NW) mk = wkey[par][lane];
          PivKey<T> gkey = mk;
          gkey.wave_reduce();
          const unsigned long long gm =
              __ballot(lane < PANEL_NW && mk.valid() && gkey.valid() && mk.row() == gkey.row());
          const int gw = gm ? __builtin_ctzll(gm) : 0;
          uint64_t *mine = p.rec + ((long)par * P + b) * REC;
          if (lane < KW)
            xstore<LOCAL>(mine + lane, tagged(tag, gkey.word(lane)));
          for (int x = lane; x < WMAX * NWD; x += 64) {
            const T v = wrow[par][gw][x / NWD];
            xstore<LOCAL>(mine + ROW0 + x, tagged(tag, word_of<T>(v, x % NWD)));
          }
          const bool own_k = (k - p.r0) / RPW == b;
          if (own_k)
            for (int x = lane; x < KREC; x += 64) {
              const T v = krow[par][x / NWD];
              xstore<LOCAL>(p.krec + (long)par * KREC + x, tagged(tag, word_of<T>(v, x % NWD)));
            }
          bool timeout = false;
          // phase 1: lane x < P reads workgroup x's key words
          PivKey<T> ck;
          ck.clear();
          if (lane < P) {
            const uint64_t *rx = p.rec + ((long)par * P + lane) * REC;
            uint64_t wd[KW];
#pragma unroll
            for (int h = 0; h < KW; ++h)
              wd[h] = __hip_atomic_load(rx + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int h = 0; h < KW; ++h) {
              long spins = 0;
              while ((wd[h] >> 32) != tag) {
                if (++spins > SPIN_LIMIT) {
                  timeout = true;
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
                wd[h] = __hip_atomic_load(rx + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
              ck.set_word(h, (uint32_t)wd[h]);
            }
          }
          PivKey<T> best = ck;
          best.wave_reduce();
          const int gpiv = best.valid() ? best.row() : k;
          // the workgroup that owns the winning row (lowest lane with the key)
          const bool is_win = lane < P && ck.valid() && best.valid() && ck.row() == gpiv;
          const unsigned long long wm = __ballot(is_win);
          const int cw = wm ? __builtin_ctzll(wm) : 0;
          // phase 2: winner row and row k
          const uint64_t *wrec = p.rec + ((long)par * P + cw) * REC + ROW0;
#pragma unroll
          for (int u = 0; u < (2 * WMAX * NWD + 63) / 64; ++u) {
            const int x = lane + 64 * u;
            if (x < 2 * WMAX * NWD) {
              const bool is_k = x >= WMAX * NWD;
              const int xx = is_k ? x - WMAX * NWD : x;
              const uint64_t *addr = is_k ? p.krec + (long)par * KREC + xx : wrec + xx;
              uint64_t wd = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              long spins = 0;
              while ((wd >> 32) != tag) {
                if (++spins > SPIN_LIMIT) {
                  timeout = true;
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
                wd = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
              stage[par][x] = (uint32_t)wd;
            }
          }
          if (__any(timeout) && lane == 0) {
            s_abort = 1;
            __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          if (lane == 0) s_piv[par] = gpiv;
        }
        __syncthreads();
        if (s_abort) return;
        piv = s_piv[par];
        win_row = reinterpret_cast<const T *>(&stage[par][0]);
        k_row = reinterpret_cast<const T *>(&stage[par][WMAX * NWD]);
      }
      // 4. pivot row and row k for everyone
      const T *pr;
      const T *dr;
      if (P == 1) {
        pr = (piv == k) ? krow[par] : wrow[par][s_gw[par]];
        dr = krow[par];
      } else {
        pr = (piv == k) ? k_row : win_row;
        dr = k_row;
      }
      if (b == 0 && t == 0) p.ipiv[k] = piv;
      if (t == 0) s_hist[j] = piv;
      // the pivot row into registers (16-byte LDS reads; every row of
      // wrow/krow/stage is 16-byte aligned): the update needs it, and row k
      // takes it over on an interchange
      // (f64 rows of 64 columns: two rows in registers at once would not
      // fit, so row k reads the pivot row from LDS before it is loaded)
      constexpr bool PRA_FIRST = !(sizeof(T) == 8 && WMAX == 64);
      T pra[WMAX];
      if constexpr (PRA_FIRST) lds_load_row<T, WMAX>(pr, pra);
      if (piv != k) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] == k) {
            if constexpr (PRA_FIRST) {
#pragma unroll
              for (int c = 0; c < WMAX; ++c) a[q][c] = pra[c];
            } else {
              lds_load_row<T, WMAX>(pr, a[q]);
            }
          } else if (row[q] == piv) {
            lds_load_row<T, WMAX>(dr, a[q]);
          }
        }
      }
      if constexpr (!PRA_FIRST) lds_load_row<T, WMAX>(pr, pra);
      // 5. scale (1/pivot, then multiply) and rank-1 update of the columns
      //    still to factor, fused with the rotation: column c lands in slot
      //    c-1 and column 0 (now L) in the last slot
      const T pv = pra[0];
      if (!((T)fabs(pv) < p.eps)) {
        const T inv = T(1) / pv;
        const int rem = w - j;  // columns 1..rem-1 are unfinished
        T lq[RPT], a0[RPT];
        bool below[RPT], all_below = true;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          below[q] = row[q] > k && row[q] < rend;
          all_below = all_below && below[q];
          const T l = a[q][0] * inv;
          a0[q] = below[q] ? l : a[q][0];
          lq[q] = l;
        }
        // finished columns (c >= rem) and rows at or above k (and padding
        // rows) keep their values exactly: selects, not branches
        if (__all(all_below)) {
#pragma unroll
          for (int c = 1; c < WMAX; ++c) {
            const bool live = c < rem;
            const T prc = pra[c];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
              const T u = fma(-prc, lq[q], a[q][c]);
              a[q][c - 1] = live ? u : a[q][c];
            }
          }
        } else {
#pragma unroll
          for (int c = 1; c < WMAX; ++c) {
            const bool live = c < rem;
            const T prc = pra[c];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
              const T u = fma(-prc, lq[q], a[q][c]);
              a[q][c - 1] = (live && below[q]) ? u : a[q][c];
            }
          }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) a[q][WMAX - 1] = a0[q];
      } else {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const T f = a[q][0];
#pragma unroll
          for (int c = 0; c + 1 < WMAX; ++c) a[q][c] = a[q][c + 1];
          a[q][WMAX - 1] = f;
        }
      }
      ++step;
    } else {
      // 6. column not pivoted (k >= n-1): rotate only
      if (t == 0) s_hist[j] = k;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const T f = a[q][0];
#pragma unroll
        for (int c = 0; c + 1 < WMAX; ++c) a[q][c] = a[q][c + 1];
        a[q][WMAX - 1] = f;
      }
    }
  }
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
  // the leaf's interchanges as row pairs, once, for every later row move
  if (b == 0) {
    __syncthreads();
    if (wave == 0) swap_pairs(lane < w ? s_hist[lane] : 0, p.c0, w, p.pairs);
  }
}

constexpr int WREC_MAX = 256;  // sizing of the exchange record buffer (64 workgroups x 4)

// ---------------------------------------------------------------------------
// panel_gepp_kernel with the rank-1 update taken off the column step's
// critical path (P > 1 workgroups).  In panel_gepp_kernel every thread
// updates all its unfinished columns before the next pivot search, and the
// exchange then waits on memory with the update already paid.  Here, once
// the pivot row is in, only the next column is updated (the eager column),
// the next search runs on it, and the full fused update and rotation of
// the step (the pending update) run while the next step's keys travel:
//   1. every wavefront finds its candidate on the eager column and parks
//      it, and row k, in LDS as they are in registers (before the pending
//      update); barrier;
//   2. wavefront 0 merges, publishes the key and the winning row and (from
//      the workgroup holding it) row k, each word brought up to date on the
//      way out: the same fma with the same operands as the register update
//      (one per word instead of one per element of every row);
//   3. every wavefront applies the pending update to its registers;
//   4. wavefront 0 polls the keys, fetches the winning row and row k;
//      barrier; swap, scale, eager column.
// The results are the same bits as panel_gepp_kernel's.
template <typename T, int WMAX, int RPT, bool LOCAL>
__global__ __launch_bounds__(PanelNT<T>::value) void panel_gepp_la_kernel(PanelP<T> p) {
  if (p.prio) __builtin_amdgcn_s_setprio(3);
  constexpr int PANEL_NT = PanelNT<T>::value;
  constexpr int PANEL_NW = PANEL_NT / 64;
  constexpr int RPW = PANEL_NT * RPT;
  constexpr int NWD = Words<T>::N;
  constexpr int KW = PivKey<T>::NW;
  constexpr int ROW0 = 4;
  constexpr int REC = ROW0 + WMAX * NWD;
  constexpr int KREC = WMAX * NWD;
  if (panel_idle<LOCAL>()) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int b = panel_wg<LOCAL>();
  const int P = p.nwg;
  const int rend = p.r0 + p.m;
  const int w = p.w;

  __shared__ __align__(16) T wrow[2][PANEL_NW][WMAX];  // candidates, not yet updated
  __shared__ __align__(16) PivKey<T> wkey[2][PANEL_NW];
  __shared__ __align__(16) T krow[2][WMAX];            // row k, not yet updated
  __shared__ __align__(16) uint32_t stage[2][2 * WMAX * NWD];  // pivot row, row k
  __shared__ int s_piv[2];
  __shared__ int s_hist[WMAX];
  __shared__ int s_abort;

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

  // the pending update: 0 none, 1 rotation only (pivot below eps), 2 scale
  // and update; its pivot row is stage[ppar] (row k's half if pk)
  int pend = 0, ppar = 0, prem = 0, pk = 0, kprev = 0;
  T pinv = T(0);
  T nxt[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) nxt[q] = a[q][0];

  auto pend_row = [&]() -> const T * {
    return reinterpret_cast<const T *>(&stage[ppar][pk ? WMAX * NWD : 0]);
  };
  // the pending update of the registers, fused with the rotation
  auto apply_pending = [&]() {
    if (pend == 2) {
      T pra[WMAX];
      lds_load_row<T, WMAX>(pend_row(), pra);
      bool below[RPT], all_below = true;
      T lq[RPT], a0[RPT];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        below[q] = row[q] > kprev && row[q] < rend;
        all_below = all_below && below[q];
        const T l = a[q][0] * pinv;
        a0[q] = below[q] ? l : a[q][0];
        lq[q] = l;
      }
      if (__all(all_below)) {
#pragma unroll
        for (int c = 1; c < WMAX; ++c) {
          const bool live = c < prem;
          const T prc = pra[c];
#pragma unroll
          for (int q = 0; q < RPT; ++q) {
            const T u = fma(-prc, lq[q], a[q][c]);
            a[q][c - 1] = live ? u : a[q][c];
          }
        }
      } else {
#pragma unroll
        for (int c = 1; c < WMAX; ++c) {
          const bool live = c < prem;
          const T prc = pra[c];
#pragma unroll
          for (int q = 0; q < RPT; ++q) {
            const T u = fma(-prc, lq[q], a[q][c]);
            a[q][c - 1] = (live && below[q]) ? u : a[q][c];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < RPT; ++q) a[q][WMAX - 1] = a0[q];
    } else if (pend == 1) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const T f = a[q][0];
#pragma unroll
        for (int c = 0; c + 1 < WMAX; ++c) a[q][c] = a[q][c + 1];
        a[q][WMAX - 1] = f;
      }
    }
    pend = 0;
  };
  // element v of a parked row (row index r) after the pending update: the
  // same operations as apply_pending on that row
  auto fixed = [&](const T *S, int v, int r) -> T {
    if (pend == 0) return S[v];
    if (pend == 1) return v < WMAX - 1 ? S[v + 1] : S[0];
    const bool below = r > kprev && r < rend;
    const T l = S[0] * pinv;
    if (v == WMAX - 1) return below ? l : S[0];
    const int c = v + 1;
    return (c < prem && below) ? fma(-pend_row()[c], l, S[c]) : S[c];
  };

  int step = 0;
  for (int j = 0; j < w; ++j) {
    const int k = p.c0 + j;
    if (k < p.kmax) {
      const int par = step & 1;
      // 1. wavefront candidate on the eager column; park it and row k as
      //    they are (the pending update is applied on the way out)
      PivKey<T> key;
      key.clear();
      int wbest = -1;
      if constexpr (sizeof(T) == 4) {
        float tv = -1.0f;
        int tq = -1;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const float av = fabsf(nxt[q]);
          if (row[q] < rend && row[q] >= k && av > tv) {  // NaN never wins
            tv = av;
            tq = q;
          }
        }
        const float mx = wave_max_f32(tv);
        if (mx >= 0.0f) {
#pragma unroll
          for (int q = RPT - 1; q >= 0; --q) {
            const unsigned long long msk = __ballot(tq == q && tv == mx);
            if (msk) wbest = p.r0 + b * RPW + q * PANEL_NT + wave * 64 + __builtin_ctzll(msk);
          }
          key.offer(mx, wbest);
        }
      } else {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (row[q] < rend && row[q] >= k) key.offer(nxt[q], row[q]);
        key.wave_reduce();
        wbest = key.valid() ? key.row() : -1;
      }
      if (lane == 0) wkey[par][wave] = key;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (row[q] == wbest) lds_store_row<T, WMAX>(wrow[par][wave], a[q]);
        if (row[q] == k) lds_store_row<T, WMAX>(krow[par], a[q]);
      }
      __syncthreads();
      const unsigned tag = p.epoch * 64u + (unsigned)step + 1u;
      bool timeout = false;
      int gpiv = k, cw = 0;
      if (wave == 0) {
        // 2. workgroup candidate; publish key, winning row, row k
        PivKey<T> mk;
        mk.clear();
        if (lane < PANEL_NW) mk = wkey[par][lane];
        PivKey<T> gkey = mk;
        gkey.wave_reduce();
        const unsigned long long gm =
            __ballot(lane < PANEL_NW && mk.valid() && gkey.valid() && mk.row() == gkey.row());
        const int gw = gm ? __builtin_ctzll(gm) : 0;
        uint64_t *mine = p.rec + ((long)par * P + b) * REC;
        if (lane < KW) xstore<LOCAL>(mine + lane, tagged(tag, gkey.word(lane)));
        if (gkey.valid()) {
          const int cr = gkey.row();
          for (int x = lane; x < WMAX * NWD; x += 64)
            xstore<LOCAL>(mine + ROW0 + x, tagged(tag, word_of<T>(fixed(wrow[par][gw], x / NWD, cr),
                                                                 x % NWD)));
        }
        if ((k - p.r0) / RPW == b)
          for (int x = lane; x < KREC; x += 64)
            xstore<LOCAL>(p.krec + (long)par * KREC + x,
                          tagged(tag, word_of<T>(fixed(krow[par], x / NWD, k), x % NWD)));
      }
      // 3. the pending update, while the keys travel
      apply_pending();
      if (wave == 0) {
        // 4. poll the P keys, then fetch the winning row and row k
        PivKey<T> ck;
        ck.clear();
        if (lane < P) {
          const uint64_t *rx = p.rec + ((long)par * P + lane) * REC;
          uint64_t wd[KW];
#pragma unroll
          for (int h = 0; h < KW; ++h)
            wd[h] = __hip_atomic_load(rx + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
          for (int h = 0; h < KW; ++h) {
            long spins = 0;
            while ((wd[h] >> 32) != tag) {
              if (++spins > SPIN_LIMIT) {
                timeout = true;
                break;
              }
              __builtin_amdgcn_s_sleep(1);
              wd[h] = __hip_atomic_load(rx + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            ck.set_word(h, (uint32_t)wd[h]);
          }
        }
        PivKey<T> best = ck;
        best.wave_reduce();
        gpiv = best.valid() ? best.row() : k;
        const bool is_win = lane < P && ck.valid() && best.valid() && ck.row() == gpiv;
        const unsigned long long wm = __ballot(is_win);
        cw = wm ? __builtin_ctzll(wm) : 0;
        const uint64_t *wrec = p.rec + ((long)par * P + cw) * REC + ROW0;
#pragma unroll
        for (int u = 0; u < (2 * WMAX * NWD + 63) / 64; ++u) {
          const int x = lane + 64 * u;
          if (x < 2 * WMAX * NWD) {
            const bool is_k = x >= WMAX * NWD;
            const int xx = is_k ? x - WMAX * NWD : x;
            if (is_k || gpiv != k) {
              const uint64_t *addr = is_k ? p.krec + (long)par * KREC + xx : wrec + xx;
              uint64_t wd = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              long spins = 0;
              while ((wd >> 32) != tag) {
                if (++spins > SPIN_LIMIT) {
                  timeout = true;
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
                wd = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
              stage[par][x] = (uint32_t)wd;
            }
          }
        }
        if (__any(timeout) && lane == 0) {
          s_abort = 1;
          __hip_atomic_store(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_piv[par] = gpiv;
      }
      __syncthreads();
      if (s_abort) return;
      const int piv = s_piv[par];
      if (b == 0 && t == 0) p.ipiv[k] = piv;
      if (t == 0) s_hist[j] = piv;
      // 5. swap; scale and the eager column; the rest becomes pending
      const T *dr = reinterpret_cast<const T *>(&stage[par][WMAX * NWD]);
      const T *pr = piv == k ? dr : reinterpret_cast<const T *>(&stage[par][0]);
      constexpr bool PRA_FIRST = !(sizeof(T) == 8 && WMAX == 64);
      T pra[WMAX];
      if constexpr (PRA_FIRST) lds_load_row<T, WMAX>(pr, pra);
      if (piv != k) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (row[q] == k) {
            if constexpr (PRA_FIRST) {
#pragma unroll
              for (int c = 0; c < WMAX; ++c) a[q][c] = pra[c];
            } else {
              lds_load_row<T, WMAX>(pr, a[q]);
            }
          } else if (row[q] == piv) {
            lds_load_row<T, WMAX>(dr, a[q]);
          }
        }
      }
      const T pv = PRA_FIRST ? pra[0] : pr[0];
      const T pr1 = PRA_FIRST ? pra[1] : pr[1];
      ppar = par;
      pk = piv == k ? 1 : 0;
      kprev = k;
      if (!((T)fabs(pv) < p.eps)) {
        const T inv = T(1) / pv;
        pinv = inv;
        prem = w - j;
        pend = 2;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const bool below = row[q] > k && row[q] < rend;
          const T l = a[q][0] * inv;
          nxt[q] = (1 < prem && below) ? fma(-pr1, l, a[q][1]) : a[q][1];
        }
      } else {
        pend = 1;
#pragma unroll
        for (int q = 0; q < RPT; ++q) nxt[q] = a[q][1];
      }
      ++step;
    } else {
      // column not pivoted (k >= n-1): the pending update, then rotate only
      apply_pending();
      if (t == 0) s_hist[j] = k;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const T f = a[q][0];
#pragma unroll
        for (int c = 0; c + 1 < WMAX; ++c) a[q][c] = a[q][c + 1];
        a[q][WMAX - 1] = f;
        nxt[q] = a[q][0];
      }
    }
  }
  apply_pending();
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
  if (b == 0) {
    __syncthreads();
    if (wave == 0) swap_pairs(lane < w ? s_hist[lane] : 0, p.c0, w, p.pairs);
  }
}

// Apply a leaf's interchanges, given as row pairs by the panel kernel, to
// the columns [cl, ch) outside the leaf's own columns: each thread moves
// one column's touched entries, all reads first, then all writes, so the
// permutation is applied in one pass.
template <typename T, int NP>
__global__ __launch_bounds__(256) void rowswap_kernel(T *W, long ld, int cl, int ch, const int *pairs,
                                                      int *perm, int w, int skip0, int prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  static_assert(NP <= PAIR_MAX, "pairs hold at most PAIR_MAX rows");
  __shared__ int dst[PAIR_MAX], src[PAIR_MAX];
  __shared__ int npair;
  const int t = threadIdx.x;
  if (t < PAIR_MAX) {
    // unused slots become (src[0] -> src[0]): the copy loops below then run
    // without branches; see the store order there
    const int np0 = pairs[0];
    const int s0 = pairs[1 + PAIR_MAX];
    dst[t] = t < np0 ? pairs[1 + t] : s0;
    src[t] = t < np0 ? pairs[1 + PAIR_MAX + t] : s0;
    if (t == 0) npair = np0;
  }
  __syncthreads();
  const int np = npair;
  if (np == 0) return;
  if (perm && blockIdx.x == 0) {
    // the row-index vector moves with the rows (libmat's Perm); up to
    // PAIR_MAX pairs span several wavefronts: all reads, barrier, all writes
    int pv = 0;
    if (t < np) pv = perm[src[t]];
    __syncthreads();
    if (t < np) perm[dst[t]] = pv;
  }
  int c = cl + blockIdx.x * 256 + t;
  if (skip0 >= 0 && c >= skip0) c += w;  // skip the leaf's own columns
  if (c >= ch) return;
  // all reads, then all writes.  Slots past np copy row src[0] onto itself;
  // they are stored first (descending order), so every real destination is
  // written after them and keeps its new value.
  T *Wc = W + c;
  const unsigned uld = (unsigned)ld;
  T v[NP];
#pragma unroll
  for (int x = 0; x < NP; ++x) v[x] = Wc[(unsigned long)(unsigned)src[x] * uld];
#pragma unroll
  for (int x = NP - 1; x >= 0; --x) Wc[(unsigned long)(unsigned)dst[x] * uld] = v[x];
}


// Exchange of the multi-workgroup panel, bits: 4 = the update off the
// critical path (panel_gepp_la_kernel) instead of panel_gepp_kernel; 2 = on
// one XCD through its L2 (launches of at most panel_local_max() workgroups
// that fit there at once).  Measured at n = 16384 (mixed solver's fp32 LU,
// plain fp64 LU): fp32 per 32-column leaf 133.7 us (0), 125.8 (2), 128.9
// (4), 128.3 (6); plain fp64 solve 519 ms (0), 512 (2), 499 (4, 6).
// Defaults: 2 for float, 4 for double; MATGPU_PANEL_EXCHANGE overrides.
template <typename T>
static int panel_exchange() {
  static const int v = [] {
    const char *e = getenv("MATGPU_PANEL_EXCHANGE");
    return e ? atoi(e) : (sizeof(T) == 4 ? 2 : 4);
  }();
  return v;
}
static int occupancy(const void *kernel, int threads) {
  int n = 0;
  MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0));
  return n;
}
static int panel_local_max() {
  static const int v = [] {
    // one workgroup per CU of the XCD: two per CU slowed the fp64 solve
    // from 512 to 650 ms
    const char *e = getenv("MATGPU_PANEL_LOCAL_MAX");
    return e ? atoi(e) : 32;
  }();
  return v;
}

template <typename T, int WMAX, int RPT>
static void launch_panel(hipStream_t s, PanelP<T> &p) {
  constexpr int PANEL_NT = PanelNT<T>::value;
  constexpr int RPW = PANEL_NT * RPT;
  p.nwg = ceil_div(p.m, RPW);
  // At most 64 workgroups of 256 threads: all resident at once on an idle
  // device (one per CU), which the exchange needs; a launch that could not
  // be resident would end in the bounded poll's timeout, not in a hang.
  KScope ks(s, "panel", 2.0 * p.m * p.w * sizeof(T));
  const bool la = p.nwg > 1 && (panel_exchange<T>() & 4);
  // the exchange needs every workgroup resident at once: on one XCD that
  // is its CUs times the workgroups one CU holds at this kernel's register use
  static const int per_cu[2] = {
      occupancy((const void *)panel_gepp_kernel<T, WMAX, RPT, true>, PANEL_NT),
      occupancy((const void *)panel_gepp_la_kernel<T, WMAX, RPT, true>, PANEL_NT)};
  const int fit = per_cu[la ? 1 : 0] * (num_cus() / NXCD);
  const bool local =
      p.nwg > 1 && (panel_exchange<T>() & 2) && p.nwg <= std::min(fit, panel_local_max());
  const dim3 grid(local ? p.nwg * NXCD : p.nwg);
  if (la && local)
    hipLaunchKernelGGL((panel_gepp_la_kernel<T, WMAX, RPT, true>), grid, dim3(PANEL_NT), 0, s, p);
  else if (la)
    hipLaunchKernelGGL((panel_gepp_la_kernel<T, WMAX, RPT, false>), grid, dim3(PANEL_NT), 0, s, p);
  else if (local)
    hipLaunchKernelGGL((panel_gepp_kernel<T, WMAX, RPT, true>), grid, dim3(PANEL_NT), 0, s, p);
  else
    hipLaunchKernelGGL((panel_gepp_kernel<T, WMAX, RPT, false>), grid, dim3(PANEL_NT), 0, s, p);
  MG_HIP(hipGetLastError());
}

// Apply the interchanges of the leaf starting at column k0 (its row pairs,
// pairs + k0 * PAIR_STRIDE) to columns [cl, ch) of the row-major matrix,
// leaving out the leaf's own columns [k0, k0+w) when they lie in the range,
// and to perm (if not null).
template <typename T>
void lu_rowswap(hipStream_t s, T *W, long ld, const int *pairs, int *perm, int k0, int w, int cl,
                int ch) {
  MG_CHECK(w <= 64, "lu_rowswap: at most 64 interchanges per call");
  const int skip0 = (k0 >= cl && k0 + w <= ch) ? k0 : -1;
  const int count = (ch - cl) - (skip0 >= 0 ? w : 0);
  if (count <= 0 && !perm) return;
  KScope ks(s, "rowswap");
  const dim3 grid(count > 0 ? ceil_div(count, 256) : 1);
  const int *pr = pairs + (long)k0 * PAIR_STRIDE;
  if (w <= 32)
    hipLaunchKernelGGL((rowswap_kernel<T, 64>), grid, dim3(256), 0, s, W, ld, cl, ch, pr, perm, w, skip0,
                       wave_prio(s));
  else
    hipLaunchKernelGGL((rowswap_kernel<T, 128>), grid, dim3(256), 0, s, W, ld, cl, ch, pr, perm, w, skip0,
                       wave_prio(s));
  MG_HIP(hipGetLastError());
}

// The interchanges of up to BLK_LEAVES consecutive leaves (leaf l starts at
// column b0 + l * leaf) applied in order to columns [cl, ch) in one launch:
// each thread owns one column, so applying leaf after leaf in its registers
// is the sequential order; only the launches are saved.
constexpr int BLK_LEAVES = 16;

template <typename T>
__global__ __launch_bounds__(256) void rowswap_block_kernel(T *W, long ld, int cl, int ch,
                                                            const int *pairs, int b0, int leaf,
                                                            int nleaf, int prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  constexpr int NP = 64;  // leaves of at most 32 columns
  __shared__ int dst[BLK_LEAVES][NP], src[BLK_LEAVES][NP];
  __shared__ int npair[BLK_LEAVES];
  const int t = threadIdx.x;
  for (int i = t; i < nleaf * NP; i += 256) {
    const int l = i / NP, x = i % NP;
    const int *pr = pairs + (long)(b0 + l * leaf) * PAIR_STRIDE;
    const int np0 = pr[0];
    const int s0 = pr[1 + PAIR_MAX];
    // unused slots copy row src[0] onto itself, stored before every real one
    dst[l][x] = x < np0 ? pr[1 + x] : s0;
    src[l][x] = x < np0 ? pr[1 + PAIR_MAX + x] : s0;
    if (x == 0) npair[l] = np0;
  }
  __syncthreads();
  const int c = cl + blockIdx.x * 256 + t;
  if (c >= ch) return;
  T *Wc = W + c;
  const unsigned uld = (unsigned)ld;
  for (int l = 0; l < nleaf; ++l) {
    if (npair[l] == 0) continue;
    T v[NP];
#pragma unroll
    for (int x = 0; x < NP; ++x) v[x] = Wc[(unsigned long)(unsigned)src[l][x] * uld];
#pragma unroll
    for (int x = NP - 1; x >= 0; --x) Wc[(unsigned long)(unsigned)dst[l][x] * uld] = v[x];
  }
}

// The interchanges of the factored block [b0, b1) (leaves of `leaf`
// columns, pairs + k * PAIR_STRIDE for the leaf at column k) applied to
// columns [cl, ch), which lie outside the block.
template <typename T>
void lu_rowswap_block(hipStream_t s, T *W, long ld, const int *pairs, int b0, int b1, int leaf,
                      int cl, int ch) {
  if (ch <= cl || b1 <= b0) return;
  const int nleaf = ceil_div(b1 - b0, leaf);
  if (leaf > 32 || nleaf > BLK_LEAVES) {
    for (int k = b0; k < b1; k += leaf)
      lu_rowswap<T>(s, W, ld, pairs, nullptr, k, std::min(leaf, b1 - k), cl, ch);
    return;
  }
  MG_CHECK(ch <= b0 || cl >= b1, "lu_rowswap_block: the columns must lie outside the block");
  KScope ks(s, "rowswap");
  hipLaunchKernelGGL((rowswap_block_kernel<T>), dim3(ceil_div(ch - cl, 256)), dim3(256), 0, s, W,
                     ld, cl, ch, pairs, b0, leaf, nleaf, wave_prio(s));
  MG_HIP(hipGetLastError());
}

// Factor the leaf block W[c0:n, c0:c0+w] (w <= 32) and apply its row swaps
// to the columns [cl, ch) of the n x n row-major matrix outside the leaf
// (and to perm).
template <typename T, int WMAX>
static void lu_leaf_w(hipStream_t s, int n, T *W, long ld, int c0, int w, int *ipiv, int *pairs,
                      T eps) {
  constexpr int REC = 4 + WMAX * Words<T>::N;
  // tags of the exchange words: distinct per launch, never 0, below 2^25
  // (tag = epoch * 64 + step + 1 stays within 31 bits); atomic so that
  // host threads factoring on different streams draw distinct values
  static std::atomic<unsigned> epoch{0};
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
  p.pairs = pairs + (long)c0 * PAIR_STRIDE;
  p.prio = wave_prio(s);
  // the exchange polls one workgroup per lane of a wavefront
  const int maxwg = 64;
  // a fresh allocation holds stale bytes, and a stale word whose high half
  // equals a launch's tag would pass as published: clear it (tag 0 is never
  // used), in stream order before the first kernel that polls it
  const size_t rec_bytes = 2L * WREC_MAX * (4 + 64 * Words<T>::N) * 8;
  const size_t krec_bytes = 2L * 64 * Words<T>::N * 8;
  bool fresh = false;
  p.rec = (uint64_t *)ws.get("panel.rec", rec_bytes, &fresh);
  if (fresh) fill_u32(s, (unsigned *)p.rec, (long)(rec_bytes / 4), 0u);
  p.krec = (uint64_t *)ws.get("panel.krec", krec_bytes, &fresh);
  if (fresh) fill_u32(s, (unsigned *)p.krec, (long)(krec_bytes / 4), 0u);
  (void)REC;
  p.err = sync_error_word(s);
  p.epoch = epoch.fetch_add(1, std::memory_order_relaxed) % ((1u << 25) - 1) + 1;
  // one row per thread where 64 workgroups suffice (a column step then
  // costs ~40% less than with two rows per thread), else two
  const int rpt = ceil_div(m, PanelNT<T>::value) > maxwg ? 2 : 1;
  const int nwg = ceil_div(m, PanelNT<T>::value * rpt);
  MG_CHECK(nwg <= maxwg, "lu_leaf: matrix too tall");
  if (rpt == 1) {
    launch_panel<T, WMAX, 1>(s, p);
  } else {
    // two f64 rows of 64 columns per thread exceed the register file
    if constexpr (sizeof(T) == 8 && WMAX == 64) MG_CHECK(false, "lu_leaf: leaf too wide for this height");
    else launch_panel<T, WMAX, 2>(s, p);
  }
}

int tslu_max_rows();
void tslu_leaf(hipStream_t s, int n, float *W, long ld, int c0, int w, int *ipiv, int *pairs,
               float eps, float *linv, int *perm, int cl, int ch, int prev_c0);

int lu_leaf_max(int m, int elem_size) {
  return (elem_size == 8 && m > 64 * PanelNT<double>::value) ? 32 : 64;
}

// Factor the leaf block W[c0:n, c0:c0+w] (w <= lu_leaf_max) and apply its row swaps
// to the columns [cl, ch) of the n x n row-major matrix outside the leaf
// (and to perm).  Leaves up to 32 columns use the 32-column kernel.
template <typename T>
void lu_leaf(hipStream_t s, int n, T *W, long ld, int c0, int w, int *ipiv, int *perm, int *pairs,
             T eps, int cl, int ch, bool tournament, T *linv, int prev_c0) {
  MG_CHECK(w >= 1 && w <= 64, "lu_leaf: leaf width must be in [1, 64]");
  bool done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (tournament && w <= 32 && n - c0 <= tslu_max_rows()) {
      // the tournament leaf moves the rows of [cl, ch) and perm itself
      tslu_leaf(s, n, W, ld, c0, w, ipiv, pairs, eps, linv, perm, cl, ch, prev_c0);
      done = true;
    }
  }
  MG_CHECK(done || (linv == nullptr && prev_c0 < 0),
           "lu_leaf: diagonal-block inverses and fused updates need the tournament leaf");
  if (!done) {
    if (w <= 32) lu_leaf_w<T, 32>(s, n, W, ld, c0, w, ipiv, pairs, eps);
    else lu_leaf_w<T, 64>(s, n, W, ld, c0, w, ipiv, pairs, eps);
    lu_rowswap<T>(s, W, ld, pairs, perm, c0, w, cl, ch);
  }
}

template void lu_leaf<float>(hipStream_t, int, float *, long, int, int, int *, int *, int *, float, int,
                             int, bool, float *, int);
template void lu_leaf<double>(hipStream_t, int, double *, long, int, int, int *, int *, int *, double,
                              int, int, bool, double *, int);
template void lu_rowswap<float>(hipStream_t, float *, long, const int *, int *, int, int, int, int);
template void lu_rowswap<double>(hipStream_t, double *, long, const int *, int *, int, int, int, int);
template void lu_rowswap_block<float>(hipStream_t, float *, long, const int *, int, int, int, int,
                                     int);
template void lu_rowswap_block<double>(hipStream_t, double *, long, const int *, int, int, int, int,
                                      int);

}  // namespace mg
