// Device helpers shared by the pivoting kernels (panel.hip, tslu.hip):
// wavefront reductions, the row-pair form of a leaf's interchanges, and the
// self-validating words of the cross-workgroup exchanges.
#pragma once
#include <cstdint>

#include "ops.hpp"

namespace mg {

// bounded waits: polls before a kernel reports a timeout instead of hanging
constexpr long SPIN_LIMIT = 1L << 24;

// A leaf's interchanges as one permutation of the touched rows, stored as
// pairs[0] = count, pairs[1 ..] = destination rows, pairs[1 + PAIR_MAX ..]
// = source rows: row dst[i] receives the original content of row src[i].
constexpr int PAIR_MAX = 128;                 // 2 * widest leaf (64)
constexpr int PAIR_STRIDE = 1 + 2 * PAIR_MAX;  // ints per leaf

// One full wavefront turns the swap sequence piv[0 .. w) (step kk swaps
// rows k0+kk and piv[kk]) into pairs.  Lane i tracks window row k0+i (the
// original row now there); lane x also tracks the x-th distinct row below
// the window.  Every index is wave-uniform, so lane reads are v_readlane.
__device__ inline void swap_pairs(int my_piv, int k0, int w, int *out) {
  const int lane = threadIdx.x & 63;
  int my_win = k0 + lane, my_outr = -1, my_outc = -1, nout = 0;
  for (int kk = 0; kk < w; ++kk) {
    const int k = k0 + kk;
    const int pr = __builtin_amdgcn_readlane(my_piv, kk);
    if (pr == k) continue;
    const int a = __builtin_amdgcn_readlane(my_win, kk);
    int bsrc;
    if (pr < k0 + w) {
      bsrc = __builtin_amdgcn_readlane(my_win, pr - k0);
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
      bsrc = __builtin_amdgcn_readlane(my_outc, slot);
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
    out[1 + at] = k0 + lane;
    out[1 + PAIR_MAX + at] = my_win;
  }
  if (opair) {
    const int at = nw + __popcll(mo & below);
    out[1 + at] = my_outr;
    out[1 + PAIR_MAX + at] = my_outc;
  }
  if (lane == 0) out[0] = nw + __popcll(mo);
}

// Wavefront max of a float (no NaN inputs): DPP within each 16-lane row,
// readlane across the four rows; every lane gets the result.
template <int CTRL>
__device__ inline float dpp_maxf(float v) {
  const float o = __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xf, 0xf, false));
  return fmaxf(v, o);
}

__device__ inline float wave_max_f32(float v) {
  v = dpp_maxf<0xB1>(v);
  v = dpp_maxf<0x4E>(v);
  v = dpp_maxf<0x141>(v);
  v = dpp_maxf<0x140>(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// A 32-bit payload with the tag of the exchange that wrote it in the high
// half: a reader accepts a word only when the tag matches, so stale words
// from earlier launches are never taken for new ones.
__device__ inline uint64_t tagged(unsigned tag, uint32_t payload) {
  return ((uint64_t)tag << 32) | payload;
}

}  // namespace mg
