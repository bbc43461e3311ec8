// fx_pcal.hip — multi-tone phase-cal extraction for gfx950 (Mode::process,
// mode.cpp:774-788; extractors pcal.cpp:333-470, 498-690, 927-1104 for real
// samples, 707-925 and 1107-1285 for complex samples).
//
// Every processed block adds its L-sample window to the band's bins: sample
// n of the window has absolute number so + n since the second boundary
// (so = datasamples + nearestsample, mode.cpp:777) and goes to bin
// (so + n) mod N, as it is (first tone at DC, or the implicit shifts) or
// times the mixer value e^{i dphi ((so + n) mod P)} (the shifting
// extractors).  Complex samples (L = the channel count) fill complex bins.
// The tones are bins of the forward N-point DFT of the summed bins.
//
//   k_pcal_accum  one workgroup per (band, block group); each thread owns the
//                 bins b = tid + k * 256 in LDS, so no two threads touch one
//                 bin and no barrier is needed between blocks; the group's
//                 sums go to a partial buffer [band][group][bins]
//   k_pcal_final  one workgroup per band: sums the groups in order, then one
//                 wavefront per tone evaluates the DFT bin and writes the
//                 tone into the result
#include <hip/hip_runtime.h>
#include "fx_kernels.h"
#include "fx_device.h"

namespace fxg {

__global__ __launch_bounds__(256) void k_pcal_accum(PcalArgs A) {
  __shared__ float acc[kPcalMaxFloats];
  const DevPcal pc = A.pc[blockIdx.x];
  const int g = blockIdx.y;
  const DevBand B = A.bands[pc.gband];
  const DevStation S = A.st[B.station];
  const DevStationSub Q = A.ss[B.station];
  const int N = pc.nbins, L = S.L, tid = threadIdx.x;
  const int per = (A.nblocks + A.ngroups - 1) / A.ngroups;
  const int blk0 = g * per, blk1 = min(A.nblocks, blk0 + per);
  const float4 lut2 = make_float4(S.lut[0], S.lut[1], S.lut[2], S.lut[3]);
  const unsigned* stream = (const unsigned*)(A.ct + Q.ctbase + (long long)B.band * Q.ctstride);
  const long long nwords = Q.ctstride >> 2;
  const unsigned mask = (1u << S.nbits) - 1u;
  const int UB = pc.csamp ? 2 * S.nbits : S.nbits;          // bits per sample of the band's stream
  const unsigned cmask = (1u << UB) - 1u;
  const float* lutc = S.nbits == 8 ? A.lut8 : A.st[B.station].lut;   // one component's levels

  for (int b = tid; b < N; b += 256) {
    if (pc.cplx) { acc[2 * b] = 0.0f; acc[2 * b + 1] = 0.0f; }
    else acc[b] = 0.0f;
  }
  for (int blk = blk0; blk < blk1; blk++) {
    const DevBlock G = A.blocks[(long long)B.station * A.nblocks + blk];
    if (G.status != 0) continue;                 // Mode::process returned before the extraction
    const long long t0 = G.nearestsample;
    const long long so = (long long)Q.datasamples + t0;
    int som = (int)(so % N);
    if (som < 0) som += N;
    int sor = pc.rot ? (int)(so % pc.period) : 0;
    if (sor < 0) sor += pc.period;
    const bool fast = G.allvalid != 0;
    // one band of 2-bit complex samples: window sample L/2 (even start) or
    // the imaginary part of sample L/2 - 1 (odd start) is erased
    const int nerase = pc.csamp && S.granularity > 1 ? L / 2 - (int)(t0 & 1) : -1;
    const bool erase_re = (t0 & 1) == 0;
    for (int b = tid; b < N; b += 256) {
      int n = b - som;
      if (n < 0) n += N;
      float ar = 0.0f, ai = 0.0f;
      if (pc.csamp) {
        for (; n < L; n += N) {
          float2 x;
          if (fast) {
            const long long bp = (t0 + n) * UB;
            const unsigned w = (bp >> 5) < nwords ? stream[bp >> 5] : 0u;
            const unsigned code = (w >> (bp & 31)) & cmask;   // a unit never straddles a word
            x = S.nbits == 8 ? make_float2(A.lut8[code & 255u], A.lut8[code >> 8])
                             : make_float2(pick4(lut2, code & 3u), pick4(lut2, code >> 2));
          } else {
            x = sample_value_c(S, Q, A.ftab, lutc, t0 + n, G.validend, B.band);
            if (n == nerase) {                 // the unpacker's erasure (k_prologue)
              x.y = 0.0f;
              if (erase_re) x.x = 0.0f;
            }
          }
          if (pc.rot) {                        // vectorMul_cf32 by the mixer
            int m = sor + n;
            if (m >= pc.period) m %= pc.period;
            const float2 r = A.aux[pc.rotoff + m];
            ar += x.x * r.x - x.y * r.y;
            ai += x.x * r.y + x.y * r.x;
          } else {
            ar += x.x;
            ai += x.y;
          }
        }
        acc[2 * b] += ar;
        acc[2 * b + 1] += ai;
        continue;
      }
      for (; n < L; n += N) {
        float x;
        if (fast) {
          const long long bp = (t0 + n) * S.nbits;
          const unsigned w = (bp >> 5) < nwords ? stream[bp >> 5] : 0u;
          const unsigned code = (w >> (bp & 31)) & mask;
          x = S.nbits == 8 ? A.lut8[code] : pick4(lut2, code);
        } else {
          x = sample_value(S, Q, A.ftab, A.lut8, lut2, t0 + n, G.validend, B.band);
        }
        if (pc.rot) {
          int m = sor + n;
          if (m >= pc.period) m %= pc.period;
          const float2 r = A.aux[pc.rotoff + m];
          ar += x * r.x;
          ai += x * r.y;
        } else {
          ar += x;
        }
      }
      if (pc.cplx) { acc[2 * b] += ar; acc[2 * b + 1] += ai; }
      else acc[b] += ar;
    }
  }
  const int nf = pc.cplx ? 2 * N : N;
  float* out = A.part + pc.partoff + (long long)g * nf;
  for (int b = tid; b < N; b += 256) {
    if (pc.cplx) { out[2 * b] = acc[2 * b]; out[2 * b + 1] = acc[2 * b + 1]; }
    else out[b] = acc[b];
  }
}

__global__ __launch_bounds__(256) void k_pcal_final(PcalArgs A) {
  __shared__ float acc[kPcalMaxFloats];
  const DevPcal pc = A.pc[blockIdx.x];
  const int N = pc.nbins, tid = threadIdx.x;
  const int nf = pc.cplx ? 2 * N : N;
  const float* part = A.part + pc.partoff;
  for (int i = tid; i < nf; i += 256) {
    float s = 0.0f;
    for (int g = 0; g < A.ngroups; g++) s += part[(long long)g * nf + i];
    acc[i] = s;
  }
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63;
  const float2* tw = A.aux + pc.twoff;
  for (int k = w; k < pc.ntones; k += 4) {
    const int idx = A.tonebin[pc.toneoff + k];
    float re = 0.0f, im = 0.0f;
    if (idx >= 0) {
      for (int b = lane; b < N; b += 64) {
        const float2 t = tw[(b * idx) % N];      // b * idx < 2^30
        if (pc.cplx) {
          const float xr = acc[2 * b], xi = acc[2 * b + 1];
          re += xr * t.x - xi * t.y;
          im += xr * t.y + xi * t.x;
        } else {
          re += acc[b] * t.x;
          im += acc[b] * t.y;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        re += __shfl_xor(re, o);
        im += __shfl_xor(im, o);
      }
    }
    if (lane == 0) {
      A.res[2 * (pc.resoff + k)] = re;
      A.res[2 * (pc.resoff + k) + 1] = im;
    }
  }
}

hipError_t launch_pcal(const PcalArgs& a, int npcal, hipStream_t stream) {
  if (npcal <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pcal_accum, dim3(npcal, a.ngroups), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_pcal_final, dim3(npcal), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace fxg
