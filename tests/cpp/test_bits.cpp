// fxg::transpose16x2 (the 2-bit corner turn of the station kernels) against
// direct field extraction on random words, and the identity of applying it twice.
#include <cstdio>
#include <cstring>
#include <random>

#include "../../difx-svn2git-try3_amd/csrc/fx_bits.h"

int main() {
  std::mt19937 rng(7);
  int bad = 0;
  for (int trial = 0; trial < 2000; trial++) {
    uint32_t in[16], w[16];
    for (int t = 0; t < 16; t++) in[t] = (uint32_t)rng();
    memcpy(w, in, sizeof w);
    fxg::transpose16x2(w);
    for (int k = 0; k < 16; k++)
      for (int t = 0; t < 16; t++)
        if (((w[k] >> (2 * t)) & 3u) != ((in[t] >> (2 * k)) & 3u)) bad++;
    fxg::transpose16x2(w);
    if (memcmp(w, in, sizeof w) != 0) bad++;
  }
  printf("%s: %d mismatches\n", bad ? "FAIL" : "PASS", bad);
  return bad ? 1 : 0;
}
