// Blocked triangular solves for a few right-hand sides (the substitutions
// of mat_solve / mat_solve_spd, mat.h:4584-4803, and every refinement step
// of the mixed-precision solvers).
//
// MI355X design: the recursion used for wide right-hand sides degenerates
// for 1..64 columns (one thread per column in the leaves).  Here the
// triangle is cut into TB x TB diagonal blocks that are inverted once per
// factorization (one workgroup per block, in LDS).  A solve is then one
// launch of a chained kernel: each row block is owned by one workgroup,
// which applies the already-solved block columns as they are published and
// finishes with its inverted diagonal block.  Every tile of the triangle is
// read once, and the only serial part is one tile product, one diagonal
// product and one flag hand-off per block.
#include <algorithm>
#include <atomic>

#include "ops.hpp"

namespace mg {

template <typename T>
struct TriTB;
template <>
struct TriTB<float> {
  static constexpr int TB = 128;
};
template <>
struct TriTB<double> {
  static constexpr int TB = 64;
};

// Stage rows [r0, r0+nr) x cols [c0, c0+nc) of a strided view into LDS
// (dst[i*ldd + j]), zero outside the range.  Loads are issued in groups of
// CH before their LDS stores, so the global latency is paid once per group.
template <typename T, int TB, int NT>
__device__ inline void stage_tile(View<const T> V, int r0, int nr, int c0, int nc, T *dst, int ldd) {
  constexpr int PER = TB * TB / NT;
  constexpr int CH = PER < 16 ? PER : 16;
  const bool rowmaj = V.cs == 1;
  for (int g = 0; g < PER; g += CH) {
    T v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int idx = threadIdx.x + (g + u) * NT;
      const int i = rowmaj ? idx / TB : idx % TB;
      const int j = rowmaj ? idx % TB : idx / TB;
      v[u] = (i < nr && j < nc) ? V(r0 + i, c0 + j) : T(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int idx = threadIdx.x + (g + u) * NT;
      const int i = rowmaj ? idx / TB : idx % TB;
      const int j = rowmaj ? idx % TB : idx / TB;
      dst[i * ldd + j] = v[u];
    }
  }
}

// Inverse of diagonal block b (rows/cols b*TB ..), written row-major to
// D + b*TB*TB; a short last block is padded with the identity.
// Divide and conquer: the 16 x 16 diagonal sub-blocks are inverted by
// substitution, one thread per column with the column in registers; then
// sizes double (16 -> 32 -> ... -> TB) with
//   [A 0; C B]^-1 = [A^-1 0; -B^-1 C A^-1  B^-1]          (lower)
//   [A C; 0 B]^-1 = [A^-1  -A^-1 C B^-1; 0 B^-1]          (upper)
// two small LDS products per level, every pair of the level at once.
template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void tri_inv_kernel(int n, View<const T> Tm, T *D) {
  constexpr int TB = TriTB<T>::TB;
  constexpr int LD = TB + 1;
  constexpr int BS = 16;
  __shared__ T S[TB * LD];       // the block
  __shared__ T X[TB * LD];       // its inverse
  __shared__ T Tmp[TB * TB / 4]; // C A^-1 (lower) / C B^-1 (upper), all pairs of a level
  const int b = blockIdx.x, o = b * TB;
  const int nr = min(TB, n - o);
  const int t = threadIdx.x;
  stage_tile<T, TB, 256>(Tm, o, nr, o, nr, S, LD);
  __syncthreads();
  if (t < TB && (t >= nr || UNIT)) S[t * LD + t] = T(1);
  for (int e = t; e < TB * TB; e += 256) X[(e / TB) * LD + e % TB] = T(0);
  __syncthreads();
  if (t < TB) {
    const int q0 = (t / BS) * BS, col = t % BS;
    T x[BS];
    if (LOWER) {
#pragma unroll
      for (int r = 0; r < BS; ++r) {
        T acc = (r == col) ? T(1) : T(0);
#pragma unroll
        for (int k = 0; k < r; ++k) acc = fma(-S[(q0 + r) * LD + q0 + k], x[k], acc);
        x[r] = UNIT ? acc : acc / S[(q0 + r) * LD + q0 + r];
      }
    } else {
#pragma unroll
      for (int r = BS - 1; r >= 0; --r) {
        T acc = (r == col) ? T(1) : T(0);
#pragma unroll
        for (int k = r + 1; k < BS; ++k) acc = fma(-S[(q0 + r) * LD + q0 + k], x[k], acc);
        x[r] = UNIT ? acc : acc / S[(q0 + r) * LD + q0 + r];
      }
    }
#pragma unroll
    for (int r = 0; r < BS; ++r) X[(q0 + r) * LD + q0 + col] = x[r];
  }
  __syncthreads();
  // each thread owns 4 x 4 micro-tiles of the off-diagonal blocks; the
  // triangular operands are zero outside their triangle (X starts at zero),
  // so the k ranges may overshoot to the micro-tile's bound.  The k loops
  // are unrolled by four so that four steps' LDS reads are in flight at once.
  constexpr int MT = 4;
  for (int sz = BS; sz < TB; sz *= 2) {
    const int tiles = sz / MT;                    // micro-tiles per block side
    const int nmt = (TB / (2 * sz)) * tiles * tiles;
    for (int e = t; e < nmt; e += 256) {
      const int p = e / (tiles * tiles), i0 = (e % (tiles * tiles)) / tiles * MT, j0 = e % tiles * MT;
      const int q = 2 * sz * p;
      T acc[MT][MT] = {};
      if (LOWER) {  // Tmp = C A^-1, A^-1(k, j) = 0 for k < j
#pragma unroll 4
        for (int k = j0; k < sz; ++k) {
          T x[MT], y[MT];
#pragma unroll
          for (int u = 0; u < MT; ++u) x[u] = S[(q + sz + i0 + u) * LD + q + k];
#pragma unroll
          for (int u = 0; u < MT; ++u) y[u] = X[(q + k) * LD + q + j0 + u];
#pragma unroll
          for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int v = 0; v < MT; ++v) acc[u][v] = fma(x[u], y[v], acc[u][v]);
        }
      } else {  // Tmp = C B^-1, B^-1(k, j) = 0 for k > j
#pragma unroll 4
        for (int k = 0; k < j0 + MT; ++k) {
          T x[MT], y[MT];
#pragma unroll
          for (int u = 0; u < MT; ++u) x[u] = S[(q + i0 + u) * LD + q + sz + k];
#pragma unroll
          for (int u = 0; u < MT; ++u) y[u] = X[(q + sz + k) * LD + q + sz + j0 + u];
#pragma unroll
          for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int v = 0; v < MT; ++v) acc[u][v] = fma(x[u], y[v], acc[u][v]);
        }
      }
      T *Tp = Tmp + p * sz * sz;
#pragma unroll
      for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int v = 0; v < MT; ++v) Tp[(i0 + u) * sz + j0 + v] = acc[u][v];
    }
    __syncthreads();
    for (int e = t; e < nmt; e += 256) {
      const int p = e / (tiles * tiles), i0 = (e % (tiles * tiles)) / tiles * MT, j0 = e % tiles * MT;
      const int q = 2 * sz * p;
      const T *Tp = Tmp + p * sz * sz;
      T acc[MT][MT] = {};
      // lower: -B^-1 Tmp, B^-1(i, k) = 0 for k > i;  upper: -A^-1 Tmp, A^-1(i, k) = 0 for k < i
      const int k0 = LOWER ? 0 : i0, k1 = LOWER ? i0 + MT : sz;
      const int xr = LOWER ? q + sz : q, xc = LOWER ? q + sz : q;
#pragma unroll 4
      for (int k = k0; k < k1; ++k) {
        T x[MT], y[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) x[u] = X[(xr + i0 + u) * LD + xc + k];
#pragma unroll
        for (int u = 0; u < MT; ++u) y[u] = Tp[k * sz + j0 + u];
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
          for (int v = 0; v < MT; ++v) acc[u][v] = fma(x[u], y[v], acc[u][v]);
      }
      const int orow = LOWER ? q + sz : q, ocol = LOWER ? q : q + sz;
#pragma unroll
      for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int v = 0; v < MT; ++v) X[(orow + i0 + u) * LD + ocol + j0 + v] = -acc[u][v];
    }
    __syncthreads();
  }
  T *Db = D + (long)b * TB * TB;
  for (int e = t; e < TB * TB; e += 256) Db[e] = X[(e / TB) * LD + e % TB];
}

// Chained block solve, one launch.  Workgroups claim row blocks in solve
// order from an atomic counter (forward: 0, 1, ...; backward: nb-1, ...).
// The workgroup owning row block r streams T(r, k) for every solved block
// k in order (tile k+1 is prefetched into registers while tile k is
// applied), reads X_k, and finally applies the inverted diagonal block,
// X_r = D_r^-1 (b_r - sum T(r,k) X_k).  Solved blocks are handed over as
// 64-bit words (launch tag << 32 | 32 payload bits) written and polled with
// agent-scope atomics, so a consumer never needs a fence or a flag: a word
// with the current tag is the value.  A workgroup only waits for blocks
// claimed before its own, by workgroups already running, so the chain
// cannot deadlock whatever the residency.  Waits are bounded: on expiry the
// workgroup sets the stream's error word and exits.
constexpr int CHAIN_NC = 16;  // right-hand sides per launch
constexpr long CHAIN_SPIN = 1L << 22;

template <typename T>
struct ChainNT;
template <>
struct ChainNT<float> {
  static constexpr int NT = 512;
};
template <>
struct ChainNT<double> {
  static constexpr int NT = 256;
};

// Tile (nr x nc at (r0, c0) of V, zero padded) into per-thread registers in
// the coalesced order of V's storage: thread t holds the elements at inner
// index t % TB (contiguous in memory) and outer index t / TB + u * NT / TB.
template <typename T, int TB, int NT>
__device__ __forceinline__ void chain_load(View<const T> V, int r0, int nr, int c0, int nc,
                                           T (&v)[TB * TB / NT]) {
  constexpr int NG = NT / TB;
  const bool rm = V.cs == 1;
  const int in = threadIdx.x % TB, out0 = threadIdx.x / TB;
  const int nin = rm ? nc : nr, nout = rm ? nr : nc;
  const long ostride = rm ? V.rs : V.cs;
  const T *base = rm ? &V(r0 + out0, c0 + in) : &V(r0 + in, c0 + out0);
#pragma unroll
  for (int u = 0; u < TB * TB / NT; ++u)
    v[u] = (in < nin && out0 + u * NG < nout) ? base[u * NG * ostride] : T(0);
}

// Registers from chain_load into S[i*(TB+1) + j]; rm as in chain_load.
template <typename T, int TB, int NT>
__device__ __forceinline__ void chain_store(bool rm, const T (&v)[TB * TB / NT], T *S) {
  constexpr int NG = NT / TB;
  const int in = threadIdx.x % TB, out0 = threadIdx.x / TB;
  T *base = rm ? S + out0 * (TB + 1) + in : S + in * (TB + 1) + out0;
  const int ostep = rm ? NG * (TB + 1) : NG;
#pragma unroll
  for (int u = 0; u < TB * TB / NT; ++u) base[u * ostep] = v[u];
}

// 32-bit payload words of a value (one for float, two for double)
template <typename T>
struct ChainWords {
  static constexpr int N = sizeof(T) / 4;
};

template <typename T>
__device__ __forceinline__ uint32_t chain_word(T v, int h) {
  if constexpr (sizeof(T) == 4) {
    return __float_as_uint(v);
  } else {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return h ? (uint32_t)(u >> 32) : (uint32_t)u;
  }
}

// FOLD: tiles are the folded T'(r,k) and the workgroup starts from D_r b_r,
// so the product with the diagonal block leaves the chain's critical path.
// fp32 tile product on MFMA for the chained solve: acc (+|-)= T(16w.., :)
// X(:, 0..15) for wavefront w's 16 rows of the 128-row block, with
// v_mfma_f32_16x16x4f32 (A: lane l gives row l%16, k l/16; B: k l/16,
// column l%16; result: register r of lane l is row 4*(l/16) + r, column
// l%16).  T is in LDS as S[row * ld + k], X as Xs[k * 16 + column].  Two
// accumulators let consecutive MFMAs overlap.
typedef float chain_f32x4 __attribute__((ext_vector_type(4)));
template <bool NEG>
__device__ __forceinline__ void chain_mfma(const float *S, int ld, const float *Xs, chain_f32x4 &acc) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float *arow = S + (16 * w + (l & 15)) * ld;
  const int g = l >> 4, c = l & 15;
  chain_f32x4 acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int st = 0; st < 32; st += 2) {
    float a0 = arow[4 * st + g], a1 = arow[4 * st + 4 + g];
    const float b0 = Xs[(4 * st + g) * CHAIN_NC + c], b1 = Xs[(4 * st + 4 + g) * CHAIN_NC + c];
    if (NEG) {
      a0 = -a0;
      a1 = -a1;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc2, 0, 0, 0);
  }
  acc += acc2;
}

// The same product in two parts, so that the tile's operand (A) is read
// from LDS while the workgroup still waits for X: chain_mfma_a reads this
// lane's 32 A values; chain_mfma_b reads all 32 B values at once, then runs
// the MFMAs back to back.
__device__ __forceinline__ void chain_mfma_a(const float *S, int ld, float (&a)[32]) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float *arow = S + (16 * w + (l & 15)) * ld;
  const int g = l >> 4;
#pragma unroll
  for (int st = 0; st < 32; ++st) a[st] = arow[4 * st + g];
}
template <bool NEG>
__device__ __forceinline__ void chain_mfma_b(const float (&a)[32], const float *Xs,
                                             chain_f32x4 &acc) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4, c = l & 15;
  float b[32];
#pragma unroll
  for (int st = 0; st < 32; ++st) b[st] = Xs[(4 * st + g) * CHAIN_NC + c];
  chain_f32x4 acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int st = 0; st < 32; st += 2) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a[st] : a[st], b[st], acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a[st + 1] : a[st + 1], b[st + 1], acc2, 0,
                                                0, 0);
  }
  acc += acc2;
}

template <typename T, bool LOWER, bool FOLD>
__global__ __launch_bounds__(ChainNT<T>::NT) void tri_chain_kernel(int n, int nrhs, View<const T> Tm,
                                                                    const T *D, const T *Tp, T *X,
                                                                    long ldx, unsigned *claim,
                                                                    uint64_t *pub, unsigned tag,
                                                                    int *err) {
  constexpr int TB = TriTB<T>::TB, NT = ChainNT<T>::NT, NC = CHAIN_NC;
  constexpr int NG = NT / TB;   // column groups
  constexpr int CPT = NC / NG;  // right-hand sides per thread
  constexpr int PER = TB * TB / NT;
  constexpr int LDS = TB + 1;
  constexpr int NW = ChainWords<T>::N;
  constexpr int BW = TB * NC * NW;  // published words per block, [c][j][h]
  // fp32 solves run their tile products on MFMA (8 wavefronts x 16 rows =
  // the 128-row block, all 16 right-hand sides); unfolded ones also the
  // final product with the inverted diagonal block
  constexpr bool MF = std::is_same<T, float>::value && NT == 512 && TB == 128 && NC == 16;
  __shared__ T S[2][TB * LDS];
  __shared__ T Xs[TB * NC];  // [row j][rhs c]
  __shared__ int s_rb, s_fail;
  const int nb = (n + TB - 1) / TB;
  const int t = threadIdx.x, i = t % TB, g = t / TB;
  const bool rm = Tm.cs == 1;
  for (;;) {
    if (t == 0) {
      s_rb = (int)atomicAdd(claim, 1u);
      s_fail = 0;
    }
    __syncthreads();
    const int rb = s_rb;
    if (rb >= nb) return;
    const int r = LOWER ? rb : nb - 1 - rb;
    const int ro = r * TB, rrows = min(TB, n - ro);
    const int cnt = rb;  // blocks solved before this one
    View<const T> Dv{D + (long)r * TB * TB, TB, 1};
    // tiles: rows tr0.., columns tcol(k).. of Tv
    const View<const T> Tv =
        FOLD ? View<const T>{Tp + fold_off(LOWER, r, n, TB), fold_ld(LOWER, r, n, TB), 1} : Tm;
    const bool trm = FOLD ? true : rm;
    const int tr0 = FOLD ? 0 : ro;
    auto tcol = [&](int k) { return FOLD ? (LOWER ? k : k - r - 1) * TB : k * TB; };
    T v[PER];
    T acc[CPT];
    chain_f32x4 macc = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (FOLD) {
      // acc = D_r b_r while the earlier blocks are still being solved
      chain_load<T, TB, NT>(Dv, 0, TB, 0, TB, v);
      chain_store<T, TB, NT>(true, v, S[1]);
#pragma unroll
      for (int u = 0; u < TB * NC / NT; ++u) {
        const int e = t + u * NT;
        const int j = e % TB, c = e / TB;
        Xs[j * NC + c] = (j < rrows && c < nrhs) ? X[(ro + j) + (long)c * ldx] : T(0);
      }
      __syncthreads();
      if constexpr (MF) {
        chain_mfma<false>(S[1], LDS, Xs, macc);
      } else {
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = T(0);
#pragma unroll 8
        for (int j = 0; j < TB; ++j) {
          const T a = S[1][i * LDS + j];
#pragma unroll
          for (int q = 0; q < CPT; ++q) acc[q] = fma(a, Xs[j * NC + g * CPT + q], acc[q]);
        }
      }
      __syncthreads();
      if (cnt > 0) {
        const int k0 = LOWER ? 0 : nb - 1;
        chain_load<T, TB, NT>(Tv, tr0, rrows, tcol(k0), min(TB, n - k0 * TB), v);
      }
    } else {
      if (cnt > 0) {
        const int k0 = LOWER ? 0 : nb - 1;
        chain_load<T, TB, NT>(Tv, tr0, rrows, tcol(k0), min(TB, n - k0 * TB), v);
      } else {
        chain_load<T, TB, NT>(Dv, 0, TB, 0, TB, v);
      }
      if constexpr (MF) {
        // b_r in the accumulator layout: register q of lane l holds row
        // 16 * wave + 4 * (l / 16) + q, column l % 16
        const int l = t & 63, wv = t >> 6, c = l & 15;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 16 * wv + 4 * (l >> 4) + q;
          macc[q] = (row < rrows && c < nrhs) ? (float)X[(ro + row) + (long)c * ldx] : 0.0f;
        }
      } else {
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const int c = g * CPT + q;
          acc[q] = (i < rrows && c < nrhs) ? X[(ro + i) + (long)c * ldx] : T(0);
        }
      }
    }
    for (int idx = 0; idx < cnt; ++idx) {
      const int k = LOWER ? idx : nb - 1 - idx;
      T *Sb = S[idx & 1];
      chain_store<T, TB, NT>(trm, v, Sb);
      if (idx + 1 < cnt) {
        const int k2 = LOWER ? idx + 1 : nb - 2 - idx;
        chain_load<T, TB, NT>(Tv, tr0, rrows, tcol(k2), min(TB, n - k2 * TB), v);
      } else if (!FOLD) {
        chain_load<T, TB, NT>(Dv, 0, TB, 0, TB, v);
      }
      // MF: the tile's A operand from LDS before X_k arrives
      float areg[MF ? 32 : 1];
      if constexpr (MF) {
        __syncthreads();
        chain_mfma_a(Sb, LDS, reinterpret_cast<float(&)[32]>(areg));
      }
      // X_k from its published words: each thread polls its first word
      // alone until it carries this launch's tag (one load a thread in
      // flight while the block is not there yet), then loads its other
      // words together and polls only those still stale, so the block costs
      // one more memory round trip after it lands instead of one per word
      const uint64_t *src = pub + (long)k * BW;
      bool timeout = false;
      constexpr int PW = TB * NC / NT * NW;  // words per thread
      uint64_t wd[PW];
      auto wptr = [&](int x) {  // word x of this thread: u = x / NW, h = x % NW
        return src + (long)(t + (x / NW) * NT) * NW + x % NW;
      };
      auto poll = [&](int x) {
        long spins = 0;
        while ((unsigned)(wd[x] >> 32) != tag) {
          if (++spins > CHAIN_SPIN) {
            timeout = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          wd[x] = __hip_atomic_load(wptr(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      };
      wd[0] = __hip_atomic_load(wptr(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      poll(0);
#pragma unroll
      for (int x = 1; x < PW; ++x)
        wd[x] = __hip_atomic_load(wptr(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int x = 1; x < PW; ++x) poll(x);
#pragma unroll
      for (int u = 0; u < TB * NC / NT; ++u) {
        const int e = t + u * NT;  // = c * TB + j
        T val;
        if constexpr (NW == 1) {
          val = __uint_as_float((uint32_t)wd[u]);
        } else {
          val = __longlong_as_double(
              (long long)(((unsigned long long)(uint32_t)wd[2 * u + 1] << 32) | (uint32_t)wd[2 * u]));
        }
        Xs[(e % TB) * NC + e / TB] = val;
      }
      if (timeout) s_fail = 1;
      __syncthreads();
      if (s_fail) {
        if (t == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      if constexpr (MF) {
        chain_mfma_b<true>(reinterpret_cast<const float(&)[32]>(areg), Xs, macc);
      } else {
#pragma unroll 8
        for (int j = 0; j < TB; ++j) {
          const T a = Sb[i * LDS + j];
#pragma unroll
          for (int q = 0; q < CPT; ++q) acc[q] = fma(-a, Xs[j * NC + g * CPT + q], acc[q]);
        }
      }
      // after the last tile the folded path publishes without touching LDS
      // (and ends with its own barrier): no barrier in between
      if (!FOLD || idx + 1 < cnt) __syncthreads();
    }
    if constexpr (MF && !FOLD) {
      // X_r = D_r^-1 acc: D_r (in v since the last tile) and acc to LDS;
      // the loop's last barrier has retired every reader of Xs and of this
      // buffer
      T *Sd = S[cnt & 1];
      chain_store<T, TB, NT>(true, v, Sd);
      {
        const int l = t & 63, wv = t >> 6, c = l & 15;
#pragma unroll
        for (int q = 0; q < 4; ++q) Xs[(16 * wv + 4 * (l >> 4) + q) * NC + c] = macc[q];
      }
      __syncthreads();
      chain_f32x4 dacc = {0.0f, 0.0f, 0.0f, 0.0f};
      chain_mfma<false>(reinterpret_cast<const float *>(Sd), LDS, reinterpret_cast<const float *>(Xs),
                        dacc);
      macc = dacc;
    }
    T out[CPT];
    if constexpr (FOLD) {
#pragma unroll
      for (int q = 0; q < CPT; ++q) out[q] = acc[q];
    } else {
      // X_r = D_r^-1 acc
      T *Sd = S[cnt & 1];
      chain_store<T, TB, NT>(true, v, Sd);
#pragma unroll
      for (int q = 0; q < CPT; ++q) Xs[i * NC + g * CPT + q] = acc[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CPT; ++q) out[q] = T(0);
#pragma unroll 8
      for (int j = 0; j < TB; ++j) {
        const T a = Sd[i * LDS + j];
#pragma unroll
        for (int q = 0; q < CPT; ++q) out[q] = fma(a, Xs[j * NC + g * CPT + q], out[q]);
      }
    }
    // publish first (the chain waits on it), then the result itself
    uint64_t *dst = pub + (long)r * BW;
    if constexpr (MF) {
      // register q of lane l: row 16*wave + 4*(l/16) + q, column l%16
      const int l = t & 63, wv = t >> 6;
      const int c = l & 15;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * wv + 4 * (l >> 4) + q;
        __hip_atomic_store(dst + (long)c * TB + row,
                           ((uint64_t)tag << 32) | chain_word<T>(macc[q], 0), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * wv + 4 * (l >> 4) + q;
        if (row < rrows && c < nrhs) X[(ro + row) + (long)c * ldx] = macc[q];
      }
      __syncthreads();  // Xs and S are reused by the next claim
      continue;
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int c = g * CPT + q;
#pragma unroll
      for (int h = 0; h < NW; ++h)
        __hip_atomic_store(dst + ((long)c * TB + i) * NW + h,
                           ((uint64_t)tag << 32) | chain_word<T>(out[q], h), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int c = g * CPT + q;
      if (i < rrows && c < nrhs) X[(ro + i) + (long)c * ldx] = out[q];
    }
    __syncthreads();  // Xs and S are reused by the next claim
  }
}

template <typename T>
int tri_block_size() {
  return TriTB<T>::TB;
}

// D must hold ceil(n/TB) * TB * TB elements.
template <typename T>
void tri_invert_blocks(hipStream_t s, bool lower, bool unit, int n, View<const T> Tm, T *D) {
  constexpr int TB = TriTB<T>::TB;
  const int nb = (n + TB - 1) / TB;
  if (nb == 0) return;
  KScope ks(s, "trsm");
  if (lower && unit) hipLaunchKernelGGL((tri_inv_kernel<T, true, true>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else if (lower) hipLaunchKernelGGL((tri_inv_kernel<T, true, false>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else if (unit) hipLaunchKernelGGL((tri_inv_kernel<T, false, true>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  else hipLaunchKernelGGL((tri_inv_kernel<T, false, false>), dim3(nb), dim3(256), 0, s, n, Tm, D);
  MG_HIP(hipGetLastError());
}

// Tags of the published words: one counter for every element type, since
// f32 and f64 solves on a stream share the publication buffer.
static std::atomic<unsigned> g_chain_launches{0};

// Elements of the folded-tile buffer of one triangle.
template <typename T>
size_t tri_fold_size(bool lower, int n) {
  constexpr int TB = TriTB<T>::TB;
  const int nb = (n + TB - 1) / TB;
  if (nb <= 1) return 1;
  return (size_t)(lower ? fold_off(true, nb, n, TB) : fold_off(false, nb - 1, n, TB));
}

// Tp := D_r T(r, k) for every off-diagonal tile: all block rows of the
// triangle in one launch of MFMA tile products.
template <typename T>
void tri_fold_tiles(hipStream_t s, bool lower, int n, View<const T> Tm, const T *D, T *Tp) {
  gemm_fold_rows<T>(s, lower, n, TriTB<T>::TB, Tm, D, Tp);
}

// X := op(T)^-1 X, X column-major n x nrhs; D from tri_invert_blocks and,
// if not null, Tp from tri_fold_tiles (then D is only used for D_r b_r).
template <typename T>
void tri_solve_blocked(hipStream_t s, bool lower, int n, int nrhs, View<const T> Tm, const T *D,
                       T *X, long ldx, const T *Tp) {
  constexpr int TB = TriTB<T>::TB;
  const int nb = (n + TB - 1) / TB;
  if (nb == 0 || nrhs <= 0) return;
  Workspace &ws = workspace(s);
  unsigned *claim = (unsigned *)ws.get("tri.claim", 4);
  bool fresh = false;
  const size_t words = (size_t)nb * TB * CHAIN_NC * ChainWords<T>::N;
  uint64_t *pub = (uint64_t *)ws.get("tri.pub", words * 8, &fresh);
  if (fresh) fill_u32(s, (unsigned *)pub, (long)words * 2, 0u);  // tag 0 is never used
  int *err = sync_error_word(s);
  // one workgroup per CU fits the LDS; more would only wait
  const int grid = std::min(nb, num_cus());
  for (int cb = 0; cb < nrhs; cb += CHAIN_NC) {
    const int nc = std::min(CHAIN_NC, nrhs - cb);
    unsigned tag = ++g_chain_launches;
    if (tag == 0) tag = ++g_chain_launches;
    fill_u32(s, claim, 1, 0u);
    KScope ks(s, "tri_chain", ((double)n * n / 2 + 2.0 * n * nc) * sizeof(T));
    T *Xc = X + (long)cb * ldx;
    const dim3 gr(grid), bl(ChainNT<T>::NT);
    if (Tp) {
      if (lower)
        hipLaunchKernelGGL((tri_chain_kernel<T, true, true>), gr, bl, 0, s, n, nc, Tm, D, Tp, Xc, ldx,
                           claim, pub, tag, err);
      else
        hipLaunchKernelGGL((tri_chain_kernel<T, false, true>), gr, bl, 0, s, n, nc, Tm, D, Tp, Xc, ldx,
                           claim, pub, tag, err);
    } else {
      if (lower)
        hipLaunchKernelGGL((tri_chain_kernel<T, true, false>), gr, bl, 0, s, n, nc, Tm, D, Tp, Xc, ldx,
                           claim, pub, tag, err);
      else
        hipLaunchKernelGGL((tri_chain_kernel<T, false, false>), gr, bl, 0, s, n, nc, Tm, D, Tp, Xc,
                           ldx, claim, pub, tag, err);
    }
    MG_HIP(hipGetLastError());
  }
}

#define MG_TRI_INST(T)                                                                           \
  template int tri_block_size<T>();                                                              \
  template void tri_invert_blocks<T>(hipStream_t, bool, bool, int, View<const T>, T *);          \
  template size_t tri_fold_size<T>(bool, int);                                                   \
  template void tri_fold_tiles<T>(hipStream_t, bool, int, View<const T>, const T *, T *);        \
  template void tri_solve_blocked<T>(hipStream_t, bool, int, int, View<const T>, const T *, T *, \
                                     long, const T *);
MG_TRI_INST(float)
MG_TRI_INST(double)

}  // namespace mg
