// fx_device.h — device helpers shared by the station kernels (fx_station.hip)
// and the phase-cal kernels (fx_pcal.hip): complex multiply and the
// general-path sample decoders (real and complex).
#pragma once
#include <hip/hip_runtime.h>
#include "fx_internal.h"

namespace fxg {

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// 2-bit level lookup by selects (a dynamically indexed local array would live
// in scratch memory)
__device__ __forceinline__ float pick4(const float4& l, unsigned c) {
  const float lo = (c & 1u) ? l.y : l.x;
  const float hi = (c & 1u) ? l.w : l.z;
  return (c & 2u) ? hi : lo;
}

// sample t of band `band` from the raw frames: zero outside the frame's valid
// byte range, past validend or past the data (mark5_unpack_with_offset,
// mark5_stream_unpacker.c:129-148)
__device__ __forceinline__ float sample_value(const DevStation& S, const DevStationSub& Q,
                                              const int2* __restrict__ ftab, const float* lut8,
                                              const float4& lut2, long long t, long long validend, int band) {
  if (t >= validend || t < 0) return 0.0f;
  long long f = t / S.framesamples;
  if (f >= Q.nframes) return 0.0f;
  int tin = (int)(t - f * S.framesamples);
  int grp = tin / S.groupsamples;
  int ib = grp * S.groupbytes;
  int2 z = ftab[Q.fbase + f];
  if (ib < z.x || ib >= z.y) return 0.0f;
  // INTERLACEDVDIF: bands the frame's EDV4 validity mask blanks
  // (blank_vdif_EDV4, format_vdif.c:4699-4775)
  if (Q.fblank && ((Q.fblank[f] >> band) & 1ull)) return 0.0f;
  int bit = ((tin - grp * S.groupsamples) * S.nchan + band) * S.nbits;
  unsigned char byte = Q.data[f * S.framebytes + S.headerbytes + ib + (bit >> 3)];
  if (S.nbits == 8) return lut8[byte];
  return pick4(lut2, (byte >> (bit & 7)) & 3u);
}

// complex sample t of band `band` (complex VDIF: the real part in the low
// nbits of the band's unit, the imaginary part above), zero as sample_value
// (mark5_unpack_complex_with_offset, format_vdif.c:2325-3054); lut is the
// level table of one component (4 or 256 entries)
__device__ __forceinline__ float2 sample_value_c(const DevStation& S, const DevStationSub& Q,
                                                 const int2* __restrict__ ftab, const float* lut,
                                                 long long t, long long validend, int band) {
  if (t >= validend || t < 0) return make_float2(0.0f, 0.0f);
  long long f = t / S.framesamples;
  if (f >= Q.nframes) return make_float2(0.0f, 0.0f);
  int tin = (int)(t - f * S.framesamples);
  int grp = tin / S.groupsamples;
  int ib = grp * S.groupbytes;
  int2 z = ftab[Q.fbase + f];
  if (ib < z.x || ib >= z.y) return make_float2(0.0f, 0.0f);
  int bit = ((tin - grp * S.groupsamples) * S.nchan + band) * S.unitbits;
  const unsigned char* p = Q.data + f * S.framebytes + S.headerbytes + ib + (bit >> 3);
  if (S.nbits == 8) return make_float2(lut[p[0]], lut[p[1]]);
  const unsigned code = (p[0] >> (bit & 7)) & 15u;
  return make_float2(lut[code & 3u], lut[code >> 2]);
}

}  // namespace fxg
