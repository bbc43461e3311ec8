// fx_xmac.hip — baseline-based (X-engine) kernels for gfx950:
//   k_xmac_f32   dense Hermitian cross-multiply-accumulate over every input
//                (datastream band) of one frequency, FP32 FMA: the plain
//                drop-in for Core::processdata's vectorAddProduct_cf32
//                (core.cpp:826-975) and Mode's autocorrelations
//                (mode.cpp:1292-1375)
//   k_weights    baseline and autocorrelation weights (core.cpp:998-1045,
//                mode.cpp:1299-1307)
//   k_scatter    copy into the core result layout (core.cpp:1781-1787,
//                1299-1366)
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include "fx_kernels.h"

namespace fxg {

// ---------------------------------------------------------------------------
// FP32 XMAC.  One workgroup = one channel tile of CT channels of one
// frequency; every thread owns one channel and one 8x8 tile (ta <= tb) of
// the input-by-input product matrix, accumulating
//     V[a][b][c] += X_a[c] * conj(X_b[c])
// over every block of the chunk.  Input spectra are staged through LDS, TB
// blocks at a time.
// ---------------------------------------------------------------------------


__device__ __forceinline__ float2 load_spec(const float2* p, long long i) { return p[i]; }
__device__ __forceinline__ float2 load_spec(const __half2* p, long long i) { return __half22float2(p[i]); }
__device__ __forceinline__ float2 load_spec(const __hip_bfloat162* p, long long i) {
  __hip_bfloat162 v = p[i];
  return make_float2(__bfloat162float(v.x), __bfloat162float(v.y));
}

template <typename TS>
__global__ __launch_bounds__(576) void k_xmac_f32(XmacArgs A) {
  const TS* spec = (const TS*)A.spec;
  extern __shared__ float2 sh[];              // [TB][G*8][CT]
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int CT = A.CT, TB = A.TB, NP = A.G * 8;
  const int c0 = blockIdx.x * CT;
  const int sb0 = blockIdx.y * A.split_blocks;               // this row's blocks
  const int nb = min(A.nblocks - sb0, A.split_blocks);
  float2* const Vout = A.V + (long long)blockIdx.y * A.vsplit;
  const int cl = tid % CT;
  const int tile = tid / CT;
  int ta = 0, tb = 0;
  {
    int t = tile, row = 0;
    while (row < A.G && t >= A.G - row) { t -= A.G - row; row++; }
    ta = row; tb = row + t;
  }
  const bool active = tile < A.ntiles;
  float2 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) acc[i][k] = make_float2(0.f, 0.f);

  for (int t0 = 0; t0 < nb; t0 += TB) {
    const int nstage = TB * NP * CT;
    for (int e = tid; e < nstage; e += nthreads) {
      int c = e % CT;
      int a = (e / CT) % NP;
      int tt = e / (CT * NP);
      float2 v = make_float2(0.f, 0.f);
      if (a < A.nin && t0 + tt < nb && c0 + c < A.nchan)
        v = load_spec(spec, A.in_off[a] + (long long)(sb0 + t0 + tt) * A.block_stride + cidx(c0 + c, A.gstride));
      sh[e] = v;
    }
    __syncthreads();
    if (active) {
      const int nt = min(TB, nb - t0);
      for (int tt = 0; tt < nt; tt++) {
        const float2* row = sh + (long long)tt * NP * CT;
        float2 xb[8];
#pragma unroll
        for (int k = 0; k < 8; k++) xb[k] = row[(tb * 8 + k) * CT + cl];
#pragma unroll
        for (int i = 0; i < 8; i++) {
          float2 xa = row[(ta * 8 + i) * CT + cl];
#pragma unroll
          for (int k = 0; k < 8; k++) {
            acc[i][k].x += xa.x * xb[k].x + xa.y * xb[k].y;
            acc[i][k].y += xa.y * xb[k].x - xa.x * xb[k].y;
          }
        }
      }
    }
    __syncthreads();
  }
  if (!active || c0 + cl >= A.nchan) return;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int a = ta * 8 + i;
    if (a >= A.nin) continue;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      int b = tb * 8 + k;
      if (b >= A.nin) continue;
      float2* p = Vout + ((long long)a * A.nin + b) * A.nchan + c0 + cl;
      if (A.accumulate) { float2 o = *p; o.x += acc[i][k].x; o.y += acc[i][k].y; *p = o; }
      else *p = acc[i][k];
    }
  }
}

// ---------------------------------------------------------------------------
// weights: per weight table entry the sum over blocks of w1*w2 (a baseline
// product) or of w1 (an autocorrelation), each added in block order; w is a
// recorded band's weight on INTERLACEDVDIF datastreams, else the block's
// data weight
// ---------------------------------------------------------------------------
__device__ __forceinline__ float band_weight(const DevStation* st, const DevStationSub* ss, const DevBlock* blocks,
                                             int nblocks, int d, int b, int k) {
  const DevBlock& g = blocks[(long long)d * nblocks + k];
  if (g.status != 0) return 0.0f;
  if (b < 0) return g.weight;
  return ss[d].pbw[(long long)k * st[d].nchan + b];
}

__global__ __launch_bounds__(256) void k_weights(const DevStation* __restrict__ st,
                                                 const DevStationSub* __restrict__ ss,
                                                 const DevBlock* __restrict__ blocks, int nblocks,
                                                 const DevWTerm* __restrict__ terms, int nterms,
                                                 float* __restrict__ wout) {
  // one workgroup per entry: the per-block terms are gathered in parallel
  // into LDS and one lane adds them in block order
  constexpr int CH = 4096;
  __shared__ float4 term4[CH / 4];
  float* term = (float*)term4;
  const int i = blockIdx.x;
  if (i >= nterms) return;
  const DevWTerm T = terms[i];
  float acc = 0.0f;
  for (int t0 = 0; t0 < nblocks; t0 += CH) {
    const int n = min(CH, nblocks - t0);
    for (int k = threadIdx.x; k < CH; k += blockDim.x) {
      float v = 0.0f;
      if (k < n) {
        const float w1 = band_weight(st, ss, blocks, nblocks, T.d1, T.b1, t0 + k);
        v = T.d2 >= 0 ? w1 * band_weight(st, ss, blocks, nblocks, T.d2, T.b2, t0 + k) : w1;
      }
      term[k] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < (n + 3) / 4; k++) {
        const float4 q = term4[k];
        acc += q.x;
        acc += q.y;
        acc += q.z;
        acc += q.w;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) wout[i] = acc;
}

// ---------------------------------------------------------------------------
// scatter: result[seg.resoff + c] = mean of V[seg.voff + c avg + i], i < avg
// (conjugated if flagged), scaled by vscale; float segments copy one weight
// ---------------------------------------------------------------------------
constexpr int kScatterPer = 4;   // channels per thread (their loads in flight together)

__global__ __launch_bounds__(256) void k_scatter(const DevSeg* __restrict__ segs, int nseg,
                                                 const float2* __restrict__ V, const float* __restrict__ w,
                                                 float vscale, float* __restrict__ res) {
  int sidx = blockIdx.y;
  if (sidx >= nseg) return;
  const DevSeg g = segs[sidx];
  const int cb = blockIdx.x * (256 * kScatterPer) + threadIdx.x;
  if (g.kind == SEG_FLOATS) {
    if (cb == 0) res[g.resoff] = w[g.voff];
    return;
  }
  if (cb >= g.nchan) return;
  float2 v[kScatterPer];
  if (g.avg == 1) {
#pragma unroll
    for (int q = 0; q < kScatterPer; q++) {
      const int c = cb + q * 256;
      v[q] = c < g.nchan ? V[g.voff + c] : make_float2(0.0f, 0.0f);
    }
  } else {
    // vectorMean_cf32 of the group: sum in channel order, then divide
#pragma unroll
    for (int q = 0; q < kScatterPer; q++) {
      const int c = cb + q * 256;
      v[q] = make_float2(0.0f, 0.0f);
      if (c >= g.nchan) continue;
      const float2* p = V + g.voff + (long long)c * g.avg;
      for (int i = 0; i < g.avg; i++) { v[q].x += p[i].x; v[q].y += p[i].y; }
      v[q].x /= (float)g.avg;
      v[q].y /= (float)g.avg;
    }
  }
  const bool al8 = (reinterpret_cast<unsigned long long>(res) & 7) == 0;
#pragma unroll
  for (int q = 0; q < kScatterPer; q++) {
    const int c = cb + q * 256;
    if (c >= g.nchan) break;
    float2 y = v[q];
    if (g.conj) y.y = -y.y;
    if (al8) {
      reinterpret_cast<float2*>(res)[g.resoff + c] = make_float2(y.x * vscale, y.y * vscale);   // one 8-byte store
    } else {
      res[2 * (g.resoff + c)] = y.x * vscale;
      res[2 * (g.resoff + c) + 1] = y.y * vscale;
    }
  }
}

hipError_t launch_xmac_f32(const XmacArgs& a, int prec, hipStream_t stream) {
  size_t shmem = (size_t)a.TB * a.G * 8 * a.CT * sizeof(float2);
  dim3 grid((unsigned)((a.nchan + a.CT - 1) / a.CT), (unsigned)((a.nblocks + a.split_blocks - 1) / a.split_blocks));
  int threads = a.CT * a.ntiles;
  threads = ((threads + 63) / 64) * 64;
  if (threads > 576) return hipErrorInvalidValue;
  if (prec == FXG_XMAC_FP32) hipLaunchKernelGGL(k_xmac_f32<float2>, grid, dim3(threads), shmem, stream, a);
  else if (prec == FXG_XMAC_FP16) hipLaunchKernelGGL(k_xmac_f32<__half2>, grid, dim3(threads), shmem, stream, a);
  else hipLaunchKernelGGL(k_xmac_f32<__hip_bfloat162>, grid, dim3(threads), shmem, stream, a);
  return hipGetLastError();
}

// V[i] (+)= sum_s P[s*vsplit + i]: the block-split partial sums, added in
// split order so that the result does not depend on scheduling
__global__ __launch_bounds__(256) void k_xmac_reduce(const float2* __restrict__ P, int nsplit, long long vsplit,
                                                     float2* __restrict__ V, long long n, int accumulate) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float2 acc = accumulate ? V[i] : make_float2(0.f, 0.f);
  for (int s = 0; s < nsplit; s++) {
    const float2 p = P[(long long)s * vsplit + i];
    acc.x += p.x;
    acc.y += p.y;
  }
  V[i] = acc;
}

hipError_t launch_xmac_reduce(const float2* P, int nsplit, long long vsplit, float2* V, long long n, int accumulate,
                              hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_xmac_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, P, nsplit, vsplit, V, n,
                     accumulate);
  return hipGetLastError();
}

// p[i] = 0 for i < n: four floats per thread where aligned
__global__ __launch_bounds__(256) void k_zero(float* __restrict__ p, long long n) {
  const long long i = ((long long)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i + 3 < n && ((uintptr_t)(p + i) & 15) == 0) {
    *(float4*)(p + i) = make_float4(0.f, 0.f, 0.f, 0.f);
  } else {
    for (long long k = i; k < n && k < i + 4; k++) p[k] = 0.0f;
  }
}

hipError_t launch_zero(float* p, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const long long blocks = (n + 1023) / 1024;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

hipError_t launch_weights(const DevStation* st, const DevStationSub* ss, const DevBlock* blocks, int nblocks,
                          const DevWTerm* terms, int nterms, float* w, hipStream_t stream) {
  if (nterms == 0) return hipSuccess;
  hipLaunchKernelGGL(k_weights, dim3(nterms), dim3(256), 0, stream, st, ss, blocks, nblocks, terms, nterms, w);
  return hipGetLastError();
}

hipError_t launch_scatter(const DevSeg* segs, int nseg, int maxlen, const float2* V, const float* w, float vscale,
                          float* res, hipStream_t stream) {
  if (nseg == 0) return hipSuccess;
  int blocks = (maxlen + 256 * kScatterPer - 1) / (256 * kScatterPer);
  if (blocks < 1) blocks = 1;
  // grid.y is limited to 65535: launch in slices
  for (int s0 = 0; s0 < nseg; s0 += 65535) {
    int n = nseg - s0 < 65535 ? nseg - s0 : 65535;
    hipLaunchKernelGGL(k_scatter, dim3(blocks, n), dim3(256), 0, stream, segs + s0, n, V, w, vscale, res);
  }
  return hipGetLastError();
}

}  // namespace fxg
