// fx_kurtosis.hip — spectral kurtosis of every recorded band (Mode's
// dumpkurtosis path):
//   k_kurt_accum  s1 += |X|^2, s2 += |X|^4 per (record, band, channel), the
//                 blocks of one F/X pass in block order (mode.cpp:1260-1277)
//   k_kurt_final  sk = (n/(n-1)) (n s2 / s1^2 - 1) - 1 and the channel sums
//                 of calculateAndAverageKurtosis (mode.cpp:1426-1478)
// A record is one Core process thread's block range (core.cpp:2083-2095):
// thread t of T owns n/T blocks, plus one for t < n mod T, in order.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include "fx_kernels.h"

namespace fxg {

// first block of record t: t (n/T) + min(t, n mod T)
__device__ __forceinline__ int kurt_start(int t, int n, int T) {
  const int q = n / T, r = n % T;
  return t * q + (t < r ? t : r);
}

// record of block b
__device__ __forceinline__ int kurt_thread(int b, int n, int T) {
  const int q = n / T, r = n % T;
  const int big = r * (q + 1);                 // blocks of the records with q + 1
  return b < big ? b / (q + 1) : r + (b - big) / q;
}

// e4m3 byte (bias 7, no infinities) as float, exactly
__device__ __forceinline__ float e4m3_value(unsigned b) {
  const unsigned e = (b >> 3) & 15u, m = b & 7u;
  const float v = e == 0 ? ldexpf((float)m, -9) : ldexpf((float)(8u + m), (int)e - 10);
  return (b & 0x80u) ? -v : v;
}

// spectrum value i of the job's precision as FP32, its scale removed
__device__ __forceinline__ float2 kurt_load(const KurtArgs& A, long long i, int g, int c) {
  if (A.prec == FXG_XMAC_FP32) return ((const float2*)A.spectra)[i];
  if (A.prec == FXG_XMAC_FP16) {
    const float2 v = __half22float2(((const __half2*)A.spectra)[i]);
    return make_float2(v.x * A.invscale, v.y * A.invscale);
  }
  if (A.prec == FXG_XMAC_FP8) {
    const unsigned w = ((const unsigned short*)A.spectra)[i];
    const int e = A.cexp[(long long)g * A.C + c];
    return make_float2(ldexpf(e4m3_value(w & 0xFFu), -e), ldexpf(e4m3_value(w >> 8), -e));
  }
  const __hip_bfloat162 v = ((const __hip_bfloat162*)A.spectra)[i];
  return make_float2(__bfloat162float(v.x) * A.invscale, __bfloat162float(v.y) * A.invscale);
}

// one thread per (band, channel): the pass's blocks in order, each added to
// the running sums of the record that owns it (vectorMagnitude_cf32, then
// squared, added, squared, added: mode.cpp:1262-1275)
__global__ __launch_bounds__(256) void k_kurt_accum(KurtArgs A) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int g = blockIdx.y;
  if (c >= A.C || g >= A.nbands) return;
  const long long base = (long long)g * A.band_stride + cidx(c, A.gstride);
  const int end = A.chunk0 + A.nchunk;
  int b = A.chunk0;
  while (b < end) {
    const int t = kurt_thread(b, A.nblocks, A.nrec);
    const int tend = min(end, kurt_start(t + 1, A.nblocks, A.nrec));
    float* p1 = A.s + ((long long)t * 2 * A.nbands + g) * A.C + c;
    float* p2 = p1 + (long long)A.nbands * A.C;
    float s1 = *p1, s2 = *p2;
    for (; b < tend; b++) {
      const float2 x = kurt_load(A, base + (long long)(b - A.chunk0) * A.bstride, g, c);
      const float m = sqrtf(x.x * x.x + x.y * x.y);
      const float m2 = m * m;
      s1 = s1 + m2;
      const float m4 = m2 * m2;
      s2 = s2 + m4;
    }
    *p1 = s1;
    *p2 = s2;
  }
}

// one workgroup per (band, record)
__global__ __launch_bounds__(256) void k_kurt_final(KurtArgs A, float* __restrict__ out, int* __restrict__ nzflag) {
#pragma clang fp contract(off)
  const int g = blockIdx.x, t = blockIdx.y;
  const int tid = threadIdx.x;
  const float* s1 = A.s + ((long long)t * 2 * A.nbands + g) * A.C;
  const float* s2 = s1 + (long long)A.nbands * A.C;
  int nz = 0;
  for (int c = tid; c < A.C; c += 256) nz |= s1[c] > 0.0f ? 1 : 0;
  nz = __syncthreads_or(nz);
  float* o = out + ((long long)t * A.nbands + g) * A.nout;
  if (tid == 0) nzflag[(long long)t * A.nbands + g] = nz;
  if (!nz) {
    for (int j = tid; j < A.nout; j += 256) o[j] = 0.0f;
    return;
  }
  const int n = kurt_start(t + 1, A.nblocks, A.nrec) - kurt_start(t, A.nblocks, A.nrec);
  const float fn = (float)n;
  const float corr = (float)((double)fn / (n - 1.0));
  const int kca = A.nout < A.C ? A.C / A.nout : 1;
  const float inv = (float)(1.0 / (double)(float)kca);
  for (int j = tid; j < A.nout; j += 256) {
    float acc = 0.0f;
    for (int k = 0; k < kca; k++) {
      const int c = j * kca + k;
      const float a = s1[c] * s1[c];
      const float bb = s2[c] * fn;
      float sk = bb / a;
      sk = sk + -1.0f;
      sk = sk * corr;
      sk = sk + -1.0f;
      if (kca > 1) sk = sk * inv;
      acc = k == 0 ? sk : acc + sk;
    }
    o[j] = acc;
  }
}

hipError_t launch_kurt_accum(const KurtArgs& a, hipStream_t stream) {
  if (a.nchunk <= 0 || a.nbands <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_kurt_accum, dim3((unsigned)((a.C + 255) / 256), (unsigned)a.nbands), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_kurt_final(const KurtArgs& a, float* out, int* nzflag, hipStream_t stream) {
  if (a.nbands <= 0 || a.nrec <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_kurt_final, dim3((unsigned)a.nbands, (unsigned)a.nrec), dim3(256), 0, stream, a, out, nzflag);
  return hipGetLastError();
}

}  // namespace fxg
