// fx_mux.hip — the corner turn of VDIF thread multiplexing for gfx950
// (stage 2 of vdifmux, libraries/vdifio/src/vdifmux.c:878-1000, with the
// data rule every corner turner of cornerturners.c implements): output
// sample-time t carries, in thread order, one unit per thread (its channels
// at t, unitbits bits), padded with zero units to a power of two.
#include <hip/hip_runtime.h>
#include "fx_kernels.h"

namespace fxg {

// one thread per 32-bit output word; threads 0..7 of the first row of each
// frame also write the header
__global__ __launch_bounds__(256) void k_vdifmux(MuxArgs A) {
  const int f = blockIdx.x;
  const int w = blockIdx.y * 256 + threadIdx.x;
  unsigned* frame = (unsigned*)(A.dest + (long long)f * A.framebytes);
  if (blockIdx.y == 0 && threadIdx.x < 8) frame[threadIdx.x] = A.headers[(long long)f * 8 + threadIdx.x];
  if (w >= A.outwords) return;
  unsigned* out = frame + 8;
  const unsigned long long m = A.mask[f];
  // an incomplete frame is zero, or (validity propagation) keeps the threads
  // it has, the missing threads' units zero
  if (m != A.goodmask && (!A.partial || m == 0ull)) { out[w] = 0u; return; }
  const long long* off = A.srcoff + (long long)f * A.nthreads;
  const int U = A.unitbits;
  const unsigned umask = U == 32 ? 0xFFFFFFFFu : ((1u << U) - 1u);
  const int per = 32 / U;
  unsigned v = 0u;
  for (int j = 0; j < per; j++) {
    const long long u = (long long)w * per + j;      // output unit
    const int k = (int)(u & (A.unitsper - 1));       // thread
    const long long t = u / A.unitsper;              // sample-time
    if (k >= A.nthreads || !((m >> k) & 1ull)) continue;
    const long long bit = t * U;
    const unsigned* p = (const unsigned*)(A.src + off[k]);
    v |= ((p[bit >> 5] >> (bit & 31)) & umask) << (j * U);
  }
  out[w] = v;
}

hipError_t launch_vdifmux(const MuxArgs& a, hipStream_t stream) {
  if (a.nframes <= 0) return hipSuccess;
  dim3 grid((unsigned)a.nframes, (unsigned)((a.outwords + 255) / 256));
  hipLaunchKernelGGL(k_vdifmux, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace fxg
