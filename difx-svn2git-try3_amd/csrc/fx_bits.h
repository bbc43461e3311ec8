// fx_bits.h — bit-level helpers shared by device code and host-side tests.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define FXG_HD __host__ __device__ __forceinline__
#else
#define FXG_HD inline
#endif

namespace fxg {

// Transpose a 16 x 16 matrix of 2-bit fields held in 16 words:
// on entry field k of w[t] (bits 2k..2k+1) is element (t, k); on exit field t
// of w[k] is that element.  Four block-swap stages (8, 4, 2, 1 fields), each
// exchanging the off-diagonal blocks of every 2x2 block of the stage.
FXG_HD void transpose16x2(uint32_t* w) {
#define FXG_SWAP(a, b, sh, m)                    \
  {                                              \
    uint32_t t_ = (((a) >> (sh)) ^ (b)) & (m);  \
    (b) ^= t_;                                   \
    (a) ^= t_ << (sh);                           \
  }
  for (int t = 0; t < 8; t++) FXG_SWAP(w[t], w[t + 8], 16, 0x0000FFFFu);
  for (int t = 0; t < 16; t++)
    if ((t & 4) == 0) FXG_SWAP(w[t], w[t + 4], 8, 0x00FF00FFu);
  for (int t = 0; t < 16; t++)
    if ((t & 2) == 0) FXG_SWAP(w[t], w[t + 2], 4, 0x0F0F0F0Fu);
  for (int t = 0; t < 16; t += 2) FXG_SWAP(w[t], w[t + 1], 2, 0x33333333u);
#undef FXG_SWAP
}

}  // namespace fxg
