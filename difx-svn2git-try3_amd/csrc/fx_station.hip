// fx_station.hip — station-based (F-engine) kernels for gfx950:
//   k_frames    per-frame validity / blanking zones   (mark5_stream.c:153-202,
//               format_vdif.c:4009-4060, blanker_mark5.c:53-165, 282-312)
//   k_prologue  FP64 block geometry + data weight      (mode.cpp:649-813,
//               mk5mode.cpp:104-177)
//   k_fstage<L> LUT unpack -> fringe rotation -> L-point FP32 Stockham FFT
//               -> sideband select -> fractional-sample rotation -> spectra
//               (mode.cpp:909-1310)
//   k_unpack    standalone Mk5Mode::unpack equivalent
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include "fx_kernels.h"
#include "fx_bits.h"
#include "fx_device.h"
#include <type_traits>

namespace fxg {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ld64(const unsigned char* p) {
  unsigned long long v = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) v |= (unsigned long long)p[i] << (8 * i);
  return v;
}
__device__ __forceinline__ unsigned ld32(const unsigned char* p) {
  return (unsigned)p[0] | ((unsigned)p[1] << 8) | ((unsigned)p[2] << 16) | ((unsigned)p[3] << 24);
}

constexpr unsigned long long kFill64 = 0x1122334411223344ULL;  // fill pattern, little endian
constexpr unsigned kFill32 = 0x11223344U;

// ---------------------------------------------------------------------------
// EDV4 validity (blank_vdif_EDV4, format_vdif.c:4699-4775): when the first
// frame of the buffer carries extended version 4 with a mask length of 1..64
// (and frame length, channels and bits that match the datastream), a frame
// whose invalid bit is clear and whose 64-bit validity mask (header bytes
// 24-31) is incomplete has the samples of every band whose mask bit
// (band / (nchan / mask length)) is clear blanked.  Returns those bands.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long edv4_blank(const DevStation& S, const unsigned char* first,
                                                         const unsigned char* fr) {
  const unsigned w4 = ld32(first + 16), w2 = ld32(first + 8), w3 = ld32(first + 12);
  const int eversion = (int)(w4 >> 24), masklength = (int)((w4 >> 16) & 0xFFu);
  if (eversion != 4 || masklength < 1 || masklength > 64) return 0ull;
  const int nchan = 1 << ((w2 >> 24) & 0x1Fu);
  const long long framelength = (long long)(w2 & 0xFFFFFFu) * 8;
  const int nbits = (int)((w3 >> 26) & 0x1Fu) + 1;
  if (nchan != S.nchan || framelength != S.framebytes || nbits != S.nbits) return 0ull;
  const int d = nchan / masklength;
  if (d < 1) return 0ull;
  const unsigned long long good = nchan >= 64 ? ~0ull : ((1ull << nchan) - 1ull);
  if ((ld32(fr) >> 31) & 1u) return 0ull;
  const unsigned long long vmask = ld64(fr + 24);
  if (vmask == good) return 0ull;
  unsigned long long blank = 0ull;
  for (int c = 0; c < nchan; c++)
    if (((vmask >> (c / d)) & 1ull) == 0ull) blank |= 1ull << c;
  return blank;
}

// ---------------------------------------------------------------------------
// k_frames: one thread per frame -> (startvalid, endvalid) payload byte range
// (and, INTERLACEDVDIF, the bands the frame's EDV4 mask blanks)
// ---------------------------------------------------------------------------
__global__ void k_frames(const DevStation* __restrict__ st, const DevStationSub* __restrict__ ss,
                         int2* __restrict__ ftab) {
  const int s = blockIdx.y;
  const DevStation S = st[s];
  const DevStationSub Q = ss[s];
  long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= Q.nframes) return;
  const unsigned char* fr = Q.data + f * S.framebytes;
  const unsigned char* pl = fr + S.headerbytes;
  int2 z = make_int2(0, kBig);
  bool valid = true;
  if (S.format == FXG_FORMAT_VDIF) {
    unsigned w0 = ld32(fr), w2 = ld32(fr + 8);
    if (w2 == 0u || ((w0 >> 31) & 1u)) valid = false;
  }
  if (!valid) {
    z = make_int2(kBig, 0);
  } else if (S.blanker == FXG_BLANKER_NONE) {
    z = make_int2(0, kBig);
  } else if (S.format == FXG_FORMAT_VDIF) {
    int nword = S.databytes / 8;
    if (ld64(pl) == kFill64 || ld64(pl + 8LL * (nword - 1)) == kFill64) z = make_int2(kBig, 0);
  } else {
    // Mark5B: one 32 KiB zone covers the whole 10000-byte payload
    int n = S.databytes / 8;
    bool startOK = ld64(pl) != kFill64, endOK = ld64(pl + 8LL * (n - 1)) != kFill64;
    if (!startOK && !endOK) z = make_int2(kBig, 0);
    else if (startOK && !endOK) {
      int i = n / 2, r = n - 1;
      for (int st2 = n / 4; st2 > 0; st2 >>= 1) {
        if (ld64(pl + 8LL * i) == kFill64) { r = i; i -= st2; } else i += st2;
      }
      r = 8 * r;
      if (ld32(pl + 4LL * (r / 4 - 1)) == kFill32) r -= 4;
      z = make_int2(0, r);
    } else if (!startOK && endOK) {
      int i = n / 2, r = n - 1;
      for (int st2 = n / 4; st2 > 0; st2 >>= 1) {
        if (ld64(pl + 8LL * i) != kFill64) { r = i; i -= st2; } else i += st2;
      }
      r = 8 * r;
      if (ld32(pl + 4LL * (r / 4 - 2)) == kFill32) r += 4;
      z = make_int2(r, kBig);
    }
  }
  // Mark5B decoders test the header's invalid flag for every group, whatever
  // the blanker (mark5_format_mark5b.c:1591-1593)
  if (S.format == FXG_FORMAT_MARK5B && (fr[S.headerbytes - 11] & 0x80)) z = make_int2(kBig, 0);
  ftab[Q.fbase + f] = z;
  if (Q.fblank) Q.fblank[f] = edv4_blank(S, Q.data, fr);
}

// number of samples in [t0, t1) of frame f (frame-local sample indices) whose
// decoder group lies inside the frame's valid byte range
__device__ __forceinline__ long long valid_in_frame(const DevStation& S, int2 z, long long t0, long long t1) {
  if (z.x >= z.y) return 0;
  // group g spans bytes [g*gb, g*gb+gb); the decoder tests the first byte
  long long g0 = ((long long)z.x + S.groupbytes - 1) / S.groupbytes;
  long long g1 = ((long long)z.y + S.groupbytes - 1) / S.groupbytes;
  long long a = max(t0, g0 * S.groupsamples), b = min(t1, g1 * S.groupsamples);
  return b > a ? b - a : 0;
}

// samples in [t0, t1) of frame f (absolute sample indices) of band b that
// decode nonzero: the valid groups' samples for 2-bit levels, those whose
// byte is not 128 for 8-bit ones ((b - 128) / 3.3 = 0)
__device__ __forceinline__ long long nonzero_in_frame(const DevStation& S, const DevStationSub& Q, int2 z, long long f,
                                                      long long t0, long long t1, int b) {
  const long long base = f * (long long)S.framesamples;
  if (S.nbits != 8) return valid_in_frame(S, z, t0 - base, t1 - base);
  long long n = 0;
  const unsigned char* pl = Q.data + f * S.framebytes + S.headerbytes;
  for (long long t = t0; t < t1; t++) {
    const int tin = (int)(t - base);
    const int ib = tin * S.groupbytes;                 // 8-bit: one sample-time per group
    if (ib < z.x || ib >= z.y) continue;
    if (pl[ib + b] != 128) n++;
  }
  return n;
}

// ---------------------------------------------------------------------------
// k_prologue: FP64 geometry of every (datastream, block).  The expression
// order is the reference's, with FMA contraction off, so nearestsample and
// integerdelay are the same integers the CPU computes.
// ---------------------------------------------------------------------------
__global__ void k_prologue(const DevStation* __restrict__ st, const DevStationSub* __restrict__ ss,
                           const int* __restrict__ flags, const int2* __restrict__ ftab,
                           int nblocks, DevBlock* __restrict__ out) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const int index = blockIdx.x * blockDim.x + threadIdx.x;
  if (index >= nblocks) return;
  const DevStation S = st[s];
  const DevStationSub Q = ss[s];
  DevBlock g;
  g.status = 0; g.nearestsample = 0; g.integerdelay = 0; g.intwall = 0; g.allvalid = 0; g.pad = 0;
  g.weight = 0.0f; g.fse = 0.0f; g.validend = 0x7fffffffffffffffLL; g.a = 0; g.b = 0; g.fracwall = 0;
  g.c = 0; g.wall = 0;
  bool ok = Q.usable != 0;
  if (ok) {
    int w = index / FXG_FLAGS_PER_INT;
    ok = w < Q.nflags && ((flags[Q.flagoff + w] >> (index % FXG_FLAGS_PER_INT)) & 1);
  }
  // INTERLACEDVDIF band weights stay zero unless the block unpacks (mode.cpp:661-668)
  if (Q.pbw)
    for (int b = 0; b < S.nchan; b++) Q.pbw[(long long)index * S.nchan + b] = 0.0f;
  if (!ok) { g.status = 1; out[(long long)s * nblocks + index] = g; return; }
  const double* ip = Q.interp;
  const int fftchannels = S.L;
  double fftcentre = index + 0.5;
  double averagedelay = ip[0] * fftcentre * fftcentre + ip[1] * fftcentre + ip[2];
  double fftstartmicrosec = index * fftchannels * S.sampletime;
  double starttime = (Q.offsec - Q.datasec) * 1000000.0 +
                     ((long long)Q.offns - (long long)Q.datans) / 1000.0 + fftstartmicrosec - averagedelay;
  int nearestsample = (int)(starttime / S.sampletime + 0.5);
  double walltimesecs = (Q.scanstart + Q.offsec) + Q.offns / 1.0e9 + fftstartmicrosec / 1.0e6;
  g.intwall = (int)walltimesecs;
  g.fracwall = walltimesecs - g.intwall;
  g.wall = walltimesecs;
  g.nearestsample = nearestsample;
  if (nearestsample < -1 ||
      ((((long long)nearestsample + fftchannels) / 1) * S.bpbnum) / S.bpbden > Q.nbytes) {
    g.status = 2; out[(long long)s * nblocks + index] = g; return;
  }
  if (nearestsample == -1) nearestsample = 0;
  g.nearestsample = nearestsample;
  // good samples in [nearestsample, nearestsample + L): the decoder's count
  // after Mk5Mode's granularity erasure
  long long t = nearestsample, tend = (long long)nearestsample + S.L;
  long long good = 0;
  long long unpackstart = (long long)nearestsample - (nearestsample % S.granularity);
  if (S.quirk8) {
    // 2..4-channel 8-bit decoders stop advancing at the first blank group:
    // everything after it in the unpack window decodes as zero
    long long f0 = unpackstart / S.framesamples;
    long long f1 = (tend - 1) / S.framesamples;
    for (long long f = f0; f <= f1; f++) {
      bool fv = f < Q.nframes && ftab[Q.fbase + f].x < ftab[Q.fbase + f].y;
      if (!fv) { g.validend = max(unpackstart, f * (long long)S.framesamples); break; }
    }
  }
  long long tcut = min(tend, g.validend);
  while (t < tcut) {
    long long f = t / S.framesamples;
    long long fe = min(tcut, (f + 1) * (long long)S.framesamples);
    if (f < Q.nframes) {
      int2 z = ftab[Q.fbase + f];
      long long base = f * (long long)S.framesamples;
      good += valid_in_frame(S, z, t - base, fe - base);
    }
    t = fe;
  }
  g.weight = (float)good / (float)S.L;
  g.allvalid = good == S.L ? 1 : 0;
  if (S.cplx && S.granularity > 1) {
    // one band of 2-bit complex samples (granularity 2): Mk5Mode::unpack
    // decodes L + 2 samples from the even sample unpackstart and erases
    // floats [0, m) and [L + m, L + 2) of the complex array viewed as floats
    // (mk5mode.cpp:121-140), m = nearestsample - unpackstart.  Its good count
    // is the valid samples of [unpackstart, unpackstart + L + 2) less the
    // valid floats erased: m = 0 takes both floats of sample unpackstart +
    // L/2 (inside the window), m = 1 the real float of sample unpackstart and
    // the imaginary one of sample unpackstart + L/2.  Blocks of such streams
    // take the general decode path, which zeroes the erased floats.
    const long long u0 = unpackstart;
    auto valid1 = [&](long long u) -> long long {
      const long long f = u / S.framesamples, base = f * (long long)S.framesamples;
      return f < Q.nframes ? valid_in_frame(S, ftab[Q.fbase + f], u - base, u - base + 1) : 0;
    };
    const long long vhalf = valid1(u0 + S.L / 2), vlast = valid1(u0 + S.L + 1);
    // good counts the window [nearestsample, nearestsample + L)
    const long long adj = nearestsample == u0 ? valid1(u0 + S.L) + vlast - 2 * vhalf : vlast - vhalf;
    g.weight = (float)(good + adj) / (float)S.L;
    g.allvalid = 0;
  }
  if (Q.fblank) {
    // INTERLACEDVDIF (mk5mode.cpp:145-168): each band loses the nonzero
    // samples that its frames' EDV4 masks blank, band weight = (good -
    // invalid) / L, and the block keeps good - sum(invalid) / nchan
    float gs = (float)good;
    long long totalinvalid = 0;
    bool blanked = false;
    for (int b = 0; b < S.nchan; b++) {
      long long inv = 0;
      for (long long u = nearestsample; u < tcut;) {
        const long long f = u / S.framesamples;
        const long long fe = min(tcut, (f + 1) * (long long)S.framesamples);
        if (f < Q.nframes && ((Q.fblank[f] >> b) & 1ull)) {
          blanked = true;
          inv += nonzero_in_frame(S, Q, ftab[Q.fbase + f], f, u, fe, b);
        }
        u = fe;
      }
      Q.pbw[(long long)index * S.nchan + b] = (gs - (float)inv) / (float)S.L;
      totalinvalid += inv;
    }
    gs -= (float)totalinvalid / (float)S.nchan;
    if (gs < 0.0f) gs = 0.0f;
    g.weight = gs / (float)S.L;
    if (blanked) g.allvalid = 0;
  }
  if (!(g.weight > 0.0f)) { g.status = 3; out[(long long)s * nblocks + index] = g; return; }
  double nearestsampletime = nearestsample * S.sampletime;
  g.fse = (float)(starttime - nearestsampletime);
  if (S.order == 0) {                        // post-F (mode.cpp:791-793)
    g.integerdelay = (int)averagedelay;
    g.c = averagedelay;
  } else if (S.order == 1) {                 // linear (mode.cpp:794-813)
    double d0 = ip[0] * index * index + ip[1] * index + ip[2];
    double d1 = ip[0] * (index + 0.5) * (index + 0.5) + ip[1] * (index + 0.5) + ip[2];
    double d2 = ip[0] * (index + 1) * (index + 1) + ip[1] * (index + 1) + ip[2];
    double a = d2 - d0;
    double b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
    int integerdelay = (int)b;
    b -= integerdelay;
    g.a = a; g.b = b; g.integerdelay = integerdelay;
  } else {                                   // quadratic (mode.cpp:814-838)
    double a = ip[0];
    double b = ip[1] + index * ip[0] * 2.0;
    double c = ip[2] + index * ip[1] + (index * index) * ip[0];
    int integerdelay = (int)c;
    c -= integerdelay;
    g.a = a; g.b = b; g.c = c; g.integerdelay = integerdelay;
  }
  out[(long long)s * nblocks + index] = g;
}

// ---------------------------------------------------------------------------
// sample decoding
// ---------------------------------------------------------------------------
// standalone decode of nsamp sample-times from `first` into out[nchan][nsamp]
__global__ void k_unpack(const DevStation* __restrict__ st, const DevStationSub* __restrict__ ss,
                         const int2* __restrict__ ftab, const float* __restrict__ lut8,
                         long long first, int nsamp, long long validend, float* __restrict__ out) {
  const DevStation S = *st;
  const DevStationSub Q = *ss;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (n >= nsamp) return;
  const float4 l2 = make_float4(S.lut[0], S.lut[1], S.lut[2], S.lut[3]);
  out[(long long)c * nsamp + n] = sample_value(S, Q, ftab, lut8, l2, first + n, validend, c);
}

// ---------------------------------------------------------------------------
// in-register DFTs (forward, e^{-2 pi i nk/R})
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

__device__ __forceinline__ void dft2(float2& a, float2& b) { float2 t = a; a = add2(t, b); b = sub2(t, b); }

__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  float2 a = add2(x0, x2), b = sub2(x0, x2), c = add2(x1, x3), d = sub2(x1, x3);
  x0 = add2(a, c); x2 = sub2(a, c);
  float2 md = mul_mi(d);
  x1 = add2(b, md); x3 = sub2(b, md);
}

template <int R> struct DFT;
template <> struct DFT<2> { __device__ static void run(float2* v) { dft2(v[0], v[1]); } };
template <> struct DFT<4> { __device__ static void run(float2* v) { dft4(v[0], v[1], v[2], v[3]); } };
template <> struct DFT<8> {
  __device__ static void run(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3); dft4(o0, o1, o2, o3);
    o1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));      // * e^{-i pi/4}
    o2 = mul_mi(o2);                                             // * e^{-i pi/2}
    o3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));     // * e^{-3i pi/4}
    v[0] = add2(e0, o0); v[4] = sub2(e0, o0);
    v[1] = add2(e1, o1); v[5] = sub2(e1, o1);
    v[2] = add2(e2, o2); v[6] = sub2(e2, o2);
    v[3] = add2(e3, o3); v[7] = sub2(e3, o3);
  }
};
// 16-point DFT from the first butterflies' outputs: sm[r] = v[r] + v[r + 8],
// df[r] = v[r] - v[r + 8] (r < 8), so that a caller can fold the products
// that form v into them
__device__ __forceinline__ void dft16_from_halves(const float2* sm, const float2* df, float2* v);

template <> struct DFT<16> {
  __device__ static void run(float2* v) {
    float2 sm[8], df[8];
#pragma unroll
    for (int r = 0; r < 8; r++) { sm[r] = add2(v[r], v[r + 8]); df[r] = sub2(v[r], v[r + 8]); }
    dft16_from_halves(sm, df, v);
  }
};

__device__ __forceinline__ void dft16_from_halves(const float2* sm, const float2* df, float2* v) {
  {
    // n = 4 n1 + n2, k = k1 + 4 k2
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
      // dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]) with its first sums given
      const float2 a = sm[n2], b = df[n2], c = sm[4 + n2], d = df[4 + n2];
      const float2 md = mul_mi(d);
      y[n2][0] = add2(a, c); y[n2][2] = sub2(a, c);
      y[n2][1] = add2(b, md); y[n2][3] = sub2(b, md);
    }
    // twiddles W16^(n2*k1)
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
    y[1][1] = cmulf(y[1][1], make_float2(c1, -s1));
    y[1][2] = cmulf(y[1][2], make_float2(h, -h));
    y[1][3] = cmulf(y[1][3], make_float2(s1, -c1));
    y[2][1] = cmulf(y[2][1], make_float2(h, -h));
    y[2][2] = mul_mi(y[2][2]);
    y[2][3] = cmulf(y[2][3], make_float2(-h, -h));
    y[3][1] = cmulf(y[3][1], make_float2(s1, -c1));
    y[3][2] = cmulf(y[3][2], make_float2(-h, -h));
    y[3][3] = cmulf(y[3][3], make_float2(-c1, s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
      float2 a = y[0][k1], b = y[1][k1], c = y[2][k1], d = y[3][k1];
      dft4(a, b, c, d);
      v[k1] = a; v[k1 + 4] = b; v[k1 + 8] = c; v[k1 + 12] = d;
    }
  }
}

// L = 2^LOG = 16^a * 2^q.  Pass 1 is radix 16 (Ns = 1, straight from the
// decoder), then NMID radix-16 passes, then a last pass of radix RL = 2^q (16
// when q = 0) that computes only the half of the bins the sideband keeps.
// Every FFT thread holds 16 points in every pass.
template <int L> struct FftShape {
  static constexpr int LOG = (L == 128) ? 7 : (L == 256) ? 8 : (L == 512) ? 9 : (L == 1024) ? 10
                           : (L == 2048) ? 11 : (L == 4096) ? 12 : (L == 8192) ? 13 : 14;
  static constexpr int Q = LOG % 4;
  static constexpr int RL = Q ? (1 << Q) : 16;                    // last-pass radix
  static constexpr int NMID = Q ? LOG / 4 - 1 : LOG / 4 - 2;      // radix-16 passes between
  static constexpr int NTF = L / 16;                              // FFT threads
  static constexpr int NT = NTF < 64 ? 64 : NTF;                  // block size
  static constexpr int PAD = L + L / 16;                          // padded LDS length
  static constexpr int NSL = L / RL;                              // Ns of the last pass
  static constexpr int TABS = fxg_rotator_table_entries(L);        // rotator sub/step/quad points
};

// twiddle e^{-2 pi i m/L} from split tables (m < L): one complex product
__device__ __forceinline__ float2 tw_of(const float2* lo, const float2* hi, int m) {
  return cmulf(hi[m >> 6], lo[m & 63]);
}

// w[r] = e^{-2 pi i r m/L} for r < 16 (r*m < L): w1, w2, w4, w8 from the
// tables and the rest as products of at most three of them
__device__ __forceinline__ void twiddles16(const float2* lo, const float2* hi, int m, float2* w) {
  w[0] = make_float2(1.f, 0.f);
  w[1] = tw_of(lo, hi, m);
  w[2] = tw_of(lo, hi, 2 * m);
  w[4] = tw_of(lo, hi, 4 * m);
  w[8] = tw_of(lo, hi, 8 * m);
  w[3] = cmulf(w[1], w[2]);
  w[5] = cmulf(w[1], w[4]);
  w[6] = cmulf(w[2], w[4]);
  w[7] = cmulf(w[3], w[4]);
#pragma unroll
  for (int r = 1; r < 8; r++) w[8 + r] = cmulf(w[r], w[8]);
}

// w[r] = e^{-2 pi i r m/L} for r < R (R = 2, 4, 8)
template <int R>
__device__ __forceinline__ void twiddlesR(const float2* lo, const float2* hi, int m, float2* w) {
  w[0] = make_float2(1.f, 0.f);
  w[1] = tw_of(lo, hi, m);
  if (R >= 4) {
    w[2] = tw_of(lo, hi, 2 * m);
    w[3] = cmulf(w[1], w[2]);
  }
  if (R >= 8) {
    w[4] = tw_of(lo, hi, 4 * m);
    w[5] = cmulf(w[1], w[4]);
    w[6] = cmulf(w[2], w[4]);
    w[7] = cmulf(w[3], w[4]);
  }
}

// LDS index with one pad slot per 16 complex values: spreads the strided
// Stockham writes of the early passes over the banks
__device__ __forceinline__ int pidx(int a) { return a + (a >> 4); }

// ---------------------------------------------------------------------------
// k_cornerturn: one workgroup per (frame, datastream).  De-interleaves the
// frame payload (sample-time major, channel minor) into one contiguous
// sample stream per band, unitbits per sample (nbits, or 2 nbits for complex
// samples), LSB first, so that the F-stage of one band reads L*unitbits/8
// contiguous bytes per FFT block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cornerturn(const DevStation* __restrict__ st,
                                                    const DevStationSub* __restrict__ ss,
                                                    unsigned char* __restrict__ ct) {
  extern __shared__ unsigned char pay[];
  const int s = blockIdx.y;
  const DevStation S = st[s];
  const DevStationSub Q = ss[s];
  const long long f = blockIdx.x;
  if (f >= Q.nframes) return;
  const unsigned char* src = Q.data + f * S.framebytes + S.headerbytes;
  if ((((unsigned long long)Q.data | (unsigned long long)S.framebytes | (unsigned long long)S.headerbytes) & 3) == 0) {
    // eight words per thread in flight before their LDS stores
    const unsigned* s32 = (const unsigned*)src;
    unsigned* p32 = (unsigned*)pay;
    const int nw = S.databytes / 4;
    for (int i0 = threadIdx.x; i0 < nw; i0 += 8 * blockDim.x) {
      unsigned w[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int i = i0 + q * blockDim.x;
        if (i < nw) w[q] = s32[i];
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int i = i0 + q * blockDim.x;
        if (i < nw) p32[i] = w[q];
      }
    }
  } else {
    for (int i = threadIdx.x; i < S.databytes; i += blockDim.x) pay[i] = src[i];
  }
  __syncthreads();
  const int nbits = S.unitbits, nchan = S.nchan, bpf = S.bpf;
  if (nbits >= 8) {
    // units of whole bytes (complex 8-bit: two bytes per unit)
    const int bpu = nbits / 8;
    for (int ob = threadIdx.x; ob < nchan * bpf; ob += blockDim.x) {
      const int c = ob / bpf, i = ob - c * bpf;
      const int u = i / bpu;
      ct[Q.ctbase + (long long)c * Q.ctstride + f * bpf + i] = pay[(u * nchan + c) * bpu + (i - u * bpu)];
    }
    return;
  }
  const int per = 8 / nbits;                   // samples per output byte
  const unsigned mask = (1u << nbits) - 1u;
  for (int ob = threadIdx.x; ob < nchan * bpf; ob += blockDim.x) {
    const int c = ob / bpf, i = ob - c * bpf;
    unsigned v = 0;
    for (int k = 0; k < per; k++) {
      const int bit = ((i * per + k) * nchan + c) * nbits;
      v |= ((pay[bit >> 3] >> (bit & 7)) & mask) << (k * nbits);
    }
    ct[Q.ctbase + (long long)c * Q.ctstride + f * bpf + i] = (unsigned char)v;
  }
}

// ---------------------------------------------------------------------------
// k_cornerturn2<W>: the same corner turn for 2-bit frames whose sample-times
// are W whole 32-bit words (16 W bands), used when every datastream has that
// shape and 4W-byte aligned frames.  A workgroup stages 4096 consecutive
// sample-times through LDS with coalesced loads; each thread then turns 16
// of them for all bands with a 2-bit transpose in registers and stores one
// coalesced word per band.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void k_cornerturn2(const DevStation* __restrict__ st,
                                                     const DevStationSub* __restrict__ ss,
                                                     unsigned char* __restrict__ ct) {
  using VW = typename std::conditional<W == 1, unsigned, typename std::conditional<W == 2, uint2, uint4>::type>::type;
  constexpr int TPW = 16 * 256;                        // sample-times per workgroup
  __shared__ VW buf[TPW + TPW / 16];                   // element i at i + i/16
  const int s = blockIdx.y;
  const DevStation S = st[s];
  const DevStationSub Q = ss[s];
  const long long spf = S.databytes / (4 * W);          // sample-times per frame
  const long long total = Q.nframes * spf;
  const long long base = (long long)blockIdx.x * TPW;
  if (base >= total) return;                            // the whole workgroup
  // frame f and sample-time r within it of this thread's sample-time, kept
  // up to date as it steps by 256 (one division per workgroup)
  const long long fw = base / spf;
  long long f = fw;
  long long r = base - fw * spf + threadIdx.x;
  while (r >= spf) { r -= spf; f++; }
  // all 16 loads in flight before the first LDS store
  VW v[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const long long tau = base + threadIdx.x + k * 256;
    if (tau < total) {
      v[k] = ((const VW*)(Q.data + f * S.framebytes + S.headerbytes))[r];
    } else {
      memset(&v[k], 0, sizeof v[k]);
    }
    r += 256;
    while (r >= spf) { r -= spf; f++; }
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int i = threadIdx.x + k * 256;
    buf[i + (i >> 4)] = v[k];
  }
  __syncthreads();
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g * 16 >= total) return;
  uint32_t w[W][16];
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const VW v = buf[threadIdx.x * 17 + t];
    const uint32_t* pv = (const uint32_t*)&v;
#pragma unroll
    for (int c = 0; c < W; c++) w[c][t] = pv[c];
  }
#pragma unroll
  for (int c = 0; c < W; c++) {
    transpose16x2(w[c]);
#pragma unroll
    for (int k = 0; k < 16; k++)
      ((uint32_t*)(ct + Q.ctbase + (long long)(16 * c + k) * Q.ctstride))[g] = w[c][k];
  }
}

// ---------------------------------------------------------------------------
// Rotator tables of one (band, block) at FFT length L, by the threads tid of
// NT: FP64 phases, FP32 residual sin/cos (mode.cpp:909-1097).  rot: the pre-F
// sub/step (and quadratic) rotators; frac: the fractional-sample sub/step.
// ---------------------------------------------------------------------------
// e^{i 2 pi t} for t in turns (|t| < 256) by the hardware sine and cosine
__device__ __forceinline__ float2 turn_rotator(float t) {
  return make_float2(__builtin_amdgcn_cosf(t), __builtin_amdgcn_sinf(t));
}
// e^{i x} for an FP32 angle x in radians: x reduced exactly to turns in FP64,
// then the hardware sine and cosine (the same value sinf / cosf give x)
__device__ __forceinline__ float2 angle_rotator(float x) {
  double t = (double)x * (1.0 / kTwoPi);
  t -= rint(t);
  return turn_rotator((float)t);
}

// C: the recorded channels, L/2 (real) or L (complex) -- a compile-time
// power of two at every call, so the divisions by it are exact products
__device__ __forceinline__ void rotator_tables(const DevBand& B, const DevStation& S, const DevBlock& G, int L,
                                               int C, int tid, int NT, bool rot, bool frac, float2* rsub,
                                               float2* rstep, float2* fsub, float2* fstep, float2* rquad) {
#pragma clang fp contract(off)
  const int s = S.s;
  const double lofreq = B.lofreq, looffset = B.looffset;
  const bool looff = looffset > 0.0 || looffset < 0.0;
  double fraclooffset = 0.0;
  bool isfraclooffset = false;
  if (looff) {
    fraclooffset = fabs(looffset) - (int)fabs(looffset);
    if (fraclooffset > kTiny) isfraclooffset = true;
    if (looffset < 0) fraclooffset = -fraclooffset;
  }
  // The entries of the four tables are dealt out over the threads as one
  // index range (sub, step, frac sub, frac step), so that no wave computes
  // more than its share of the sin/cos.
  const int nsub = (rot && S.order != 0) ? s : 0;
  const int nstep = (rot && S.order != 0) ? S.nfrsteps : 0;
  const int nfsub = frac ? s : 0;
  const int nfstep = frac ? S.nfracsteps : 0;
  const int e1 = nsub, e2 = e1 + nstep, e3 = e2 + nfsub, e4 = e3 + nfstep;
  const double sign = B.useB ? -1.0 : 1.0;
  const float val = (float)((double)G.fse - B.clock + sign * B.clockdelta / 2);
  // order 0: the fringe phase of the block at its centre delay, added to
  // the sub arguments (mode.cpp:1064-1073, 1111-1119)
  float postf = 0.0f;
  if (S.order == 0 && nfsub > 0) {
    double phaserotation = (G.c - G.integerdelay) * lofreq;
    if (S.fractionalLo) phaserotation += G.integerdelay * (lofreq - (int)lofreq);
    phaserotation -= G.wall * looffset;
    postf = (float)(-kTwoPi * (phaserotation - (int)phaserotation));
  }
  // pre-F rotator points: order 1 x = a k/L + b (mode.cpp:804-812), order 2
  // x = (k/L)(b + a s/L) + c with quad points (k/L) 2a s/L and steps
  // (js/L) b + (js/L)^2 a (mode.cpp:821-837)
  const double stepxoff1 = (double)(1 * s) / (double)L;
  for (int e = tid; e < e4; e += NT) {
    if (e < e1) {
      const int k = e;
      const double subxoff = (double)k / (double)L;
      double ph = (S.order == 1 ? subxoff * G.a + G.b : subxoff * (G.b + G.a * stepxoff1) + G.c) * lofreq;
      if (S.fractionalLo) ph += (lofreq - (int)lofreq) * (double)G.integerdelay;
      if (looff) {
        ph += ((k * S.sampletime) / 1e6) * (-looffset);
        ph += -looffset * G.fracwall;
        if (isfraclooffset) ph += -fraclooffset * G.intwall;
      }
      // phase -2 pi frac(ph) (mode.cpp:962-967), as turns
      rsub[k] = turn_rotator((float)((double)(int)ph - ph));
      if (S.order == 2) {
        const double qph = (subxoff * (2 * G.a * stepxoff1)) * lofreq;
        rquad[k] = turn_rotator((float)((double)(int)qph - qph));
      }
    } else if (e < e2) {
      const int j = e - e1;
      const double stepxoff = ((double)(j * s)) / (double)L;
      double ph = (S.order == 1 ? stepxoff * G.a : stepxoff * G.b + (stepxoff * stepxoff) * G.a) * lofreq;
      if (looff) ph += (((j * s) * S.sampletime) / 1e6) * (-looffset);
      rstep[j] = turn_rotator((float)((double)(int)ph - ph));
    } else if (e < e3) {
      const int k = e - e2;
      float cf = (float)((kTwoPi * k * S.bw) / C);
      float arg = cf * val;
      if (S.order == 0) arg = arg + postf;
      // the band's polarisation phase correction rides on every channel's
      // fractional-sample rotator (a constant factor)
      fsub[k] = cmulf(angle_rotator(arg), make_float2(B.polre, B.polim));
    } else {
      const int j = e - e3;
      // step channel frequencies (mode.cpp:328-334, 864-880): DSB, LSB, USB
      double dcf = B.dsb ? (kTwoPi * j * s * S.bw) / C - kTwoPi * S.bw / 2.0
                 : B.lsb ? (-kTwoPi * ((S.nfracsteps - j) * s) * S.bw) / C : (kTwoPi * j * s * S.bw) / C;
      float cf = (float)dcf;
      fstep[j] = angle_rotator(cf * val);
    }
  }
}

// fp8 e4m3 pair (re low byte, im high byte) of y * 2^e, saturated at the
// largest finite value 448
__device__ __forceinline__ unsigned short fp8_pair(float2 y, int e) {
  const float a = fminf(fmaxf(ldexpf(y.x, e), -448.0f), 448.0f);
  const float b = fminf(fmaxf(ldexpf(y.y, e), -448.0f), 448.0f);
  return (unsigned short)(__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xFFFF);
}
struct Fp8Out {
  unsigned short* p;          // the band's spectrum of this block
  const signed char* e;       // the band's channel exponents
};

// channel c's value y stored at element i of the spectrum (i = cidx(c))
__device__ __forceinline__ void put_spec(float2* o, long long i, int c, float2 y) { o[i] = y; }
__device__ __forceinline__ void put_spec(__half2* o, long long i, int c, float2 y) { o[i] = __floats2half2_rn(y.x, y.y); }
__device__ __forceinline__ void put_spec(__hip_bfloat162* o, long long i, int c, float2 y) { o[i] = __float22bfloat162_rn(y); }
__device__ __forceinline__ void put_spec(const Fp8Out& o, long long i, int c, float2 y) { o.p[i] = fp8_pair(y, o.e[c]); }

// zero spectrum of one (band, block): C values of the job's spectrum type
__device__ __forceinline__ void zero_spec(void* spectra, int prec, long long outbase, int C, long long gs, int tid,
                                          int nt) {
  if (prec == FXG_XMAC_FP32) {
    float2* o = (float2*)spectra + outbase;
    for (int c = tid; c < C; c += nt) o[cidx(c, gs)] = make_float2(0.f, 0.f);
  } else if (prec == FXG_XMAC_FP8) {
    unsigned short* o = (unsigned short*)spectra + outbase;
    for (int c = tid; c < C; c += nt) o[cidx(c, gs)] = 0;
  } else {
    unsigned* o = (unsigned*)spectra + outbase;
    for (int c = tid; c < C; c += nt) o[cidx(c, gs)] = 0u;
  }
}

// The block-independent LDS tables of a single-workgroup F stage (split
// twiddles, the second pass's 256 twiddles, the 8-bit level table), loaded
// into registers at the start of the kernel: their loads then overlap the
// block's record loads and stream-word loads instead of each waiting on its
// own round trip before its LDS store.
template <int L, int NT, int NBITS, bool MID>
struct TabStage {
  static constexpr int NLO = (64 + NT - 1) / NT, NHI = (L / 64 + NT - 1) / NT;
  static constexpr int N256 = MID ? (256 + NT - 1) / NT : 0, NL8 = NBITS == 8 ? (256 + NT - 1) / NT : 0;
  float2 lo[NLO], hi[NHI], t256[N256 > 0 ? N256 : 1];
  float l8[NL8 > 0 ? NL8 : 1];
  __device__ __forceinline__ void load(const float2* __restrict__ tw, const float* __restrict__ lut8, int tid) {
#pragma unroll
    for (int q = 0; q < NLO; q++) { const int i = tid + q * NT; if (i < 64) lo[q] = tw[i]; }
#pragma unroll
    for (int q = 0; q < NHI; q++) { const int i = tid + q * NT; if (i < L / 64) hi[q] = tw[64 + i]; }
#pragma unroll
    for (int q = 0; q < N256; q++) { const int i = tid + q * NT; if (i < 256) t256[q] = tw[64 + L / 64 + i]; }
#pragma unroll
    for (int q = 0; q < NL8; q++) { const int i = tid + q * NT; if (i < 256) l8[q] = lut8[i]; }
  }
  __device__ __forceinline__ void store(float2* twlo, float2* twhi, float2* tw256, float* lut8s, int tid) const {
#pragma unroll
    for (int q = 0; q < NLO; q++) { const int i = tid + q * NT; if (i < 64) twlo[i] = lo[q]; }
#pragma unroll
    for (int q = 0; q < NHI; q++) { const int i = tid + q * NT; if (i < L / 64) twhi[i] = hi[q]; }
#pragma unroll
    for (int q = 0; q < N256; q++) { const int i = tid + q * NT; if (i < 256) tw256[i] = t256[q]; }
#pragma unroll
    for (int q = 0; q < NL8; q++) { const int i = tid + q * NT; if (i < 256) lut8s[i] = l8[q]; }
  }
};

// Last pass of radix 2 (L = 2 * 16^a) for thread j: butterflies k = j + u L/16,
// u < 8, with W_L^k = W_L^j W_16^u; only the kept sideband's output of each
// butterfly; channel c = k has c mod s = j mod s (one fsub value per thread,
// the power-of-two spectrum scale folded into it exactly); then the
// fractional-sample rotation fsub[c mod s] * fstep[c / s] (mode.cpp:1228-1245).
template <int L, bool LSB, typename TO>
__device__ __forceinline__ void last_pass_r2(const float2* buf, const float2* twlo, const float2* twhi,
                                             const float2* fsub, const float2* fstep, int s, int slog,
                                             float scale, TO out, long long gs, int j) {
  constexpr int NTF = L / 16, NSL = L / 2, RSTR = NTF + NTF / 16;
  // channel j + u NTF is element cj + u ustep of the spectrum (NTF is a multiple of 16)
  const long long cj = cidx(j, gs), ustep = (NTF / 16) * gs;
  const float2 wj = tw_of(twlo, twhi, j);
  float2 f0 = fsub[j & (s - 1)];
  f0 = make_float2(f0.x * scale, f0.y * scale);
  const float2* fsp = fstep + (j >> slog);
  const int fss = NTF >> slog;
  const float2* ip = buf + pidx(j);
  constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const float2 x0 = ip[u * RSTR];
    float2 t = cmulf(ip[u * RSTR + NSL + NSL / 16], wj);
    if (u == 1) t = cmulf(t, make_float2(c1, -s1));          // * W_16^u
    if (u == 2) t = cmulf(t, make_float2(h, -h));
    if (u == 3) t = cmulf(t, make_float2(s1, -c1));
    if (u == 4) t = mul_mi(t);
    if (u == 5) t = cmulf(t, make_float2(-s1, -c1));
    if (u == 6) t = cmulf(t, make_float2(-h, -h));
    if (u == 7) t = cmulf(t, make_float2(-c1, -s1));
    const float2 x = LSB ? sub2(x0, t) : add2(x0, t);
    const float2 fr = cmulf(f0, *fsp);
    fsp += fss;
    put_spec(out, cj + u * ustep, j + u * NTF, make_float2(x.x * fr.x - x.y * fr.y, x.x * fr.y + x.y * fr.x));
  }
}

template <int L, int NBITS>
__global__ __launch_bounds__(FftShape<L>::NT) void k_fstage(FStageArgs A) {
  using Sh = FftShape<L>;
  constexpr int NT = Sh::NT, NTF = Sh::NTF, RL = Sh::RL, NSL = Sh::NSL;
  constexpr int RAWW = (L * NBITS) / 32 + 2;    // staged stream words of one block
  // 8-bit samples at the largest L: the staged stream words share the tail of
  // the FFT buffer (pass 1 reads them all before its first write, below)
  constexpr bool kAliasRaw = NBITS == 8 && L >= 16384;
  // One LDS block with the small tables first, so that they lie inside the
  // 64 KiB an LDS instruction's immediate offset reaches.
  struct Smem {
    float2 tabs[Sh::TABS];                      // rotator sub/step tables, packed
    float2 twlo[64], twhi[L / 64];              // e^{-2 pi i m/L}, e^{-2 pi i 64h/L}
    float2 tw256[Sh::NMID > 0 ? 256 : 1];       // e^{-2 pi i m/256}: the second pass's twiddles
    float lut8s[NBITS == 8 ? 256 : 4];          // level table indexed by the sample code
    unsigned raw_own[kAliasRaw ? 2 : RAWW + (RAWW & 1)];
    float2 buf[Sh::PAD];
  };
  __shared__ Smem sm;
  float2 (&buf)[Sh::PAD] = sm.buf;
  float2 (&tabs)[Sh::TABS] = sm.tabs;
  float2 (&twlo)[64] = sm.twlo;
  float2 (&twhi)[L / 64] = sm.twhi;
  float2* tw256 = sm.tw256;
  float (&lut8s)[NBITS == 8 ? 256 : 4] = sm.lut8s;
  unsigned* raw = kAliasRaw ? (unsigned*)(buf + Sh::PAD) - RAWW : sm.raw_own;

  const int tid = threadIdx.x;
  // block within the chunk: workgroups are dealt to the 8 XCDs round-robin,
  // so consecutive blocks go to one XCD (its L2 merges the halves of the
  // lines that neighbouring blocks write in the grouped spectra layout)
  const int cb = (gridDim.x % 8 == 0) ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
  const int blk = A.chunk0 + cb;
  TabStage<L, NT, NBITS, (Sh::NMID > 0)> ts;
  ts.load(A.tw, A.lut8, tid);
  const int gband = A.bandlist ? A.bandlist[blockIdx.y] : blockIdx.y;
  const DevBand B = A.bands[gband];
  const DevStation S = A.st[B.station];
  const DevStationSub Q = A.ss[B.station];
  const DevBlock G = A.blocks[(long long)B.station * A.nblocks_subint + blk];
  const int C = L / 2;
  const long long outbase = B.specoff + (long long)cb * A.bstride;

  if (G.status != 0) {
    zero_spec(A.spectra, A.prec, outbase, C, A.gstride, tid, NT);
    return;
  }

  const long long t0 = G.nearestsample;
  const bool fast = G.allvalid != 0;
  // ---- stage the block's samples of this band (corner-turned stream)
  long long bit0 = t0 * NBITS;
  constexpr int NRAW = (RAWW + NT - 1) / NT;
  unsigned rw[NRAW];
  if (fast) {                                   // loads now, LDS stores after the rotator tables
    const long long w0 = bit0 >> 5;
    const unsigned* src = (const unsigned*)(A.ct + Q.ctbase + (long long)B.band * Q.ctstride);
    const long long nwords = Q.ctstride >> 2;
#pragma unroll
    for (int q = 0; q < NRAW; q++) {
      const int i = tid + q * NT;
      rw[q] = (i < RAWW && w0 + i < nwords) ? src[w0 + i] : 0u;
    }
  }

  float2* rsub = tabs;
  float2* rstep = tabs + S.s;
  float2* fsub = rstep + S.nfrsteps;
  float2* fstep = fsub + S.s;
  float2* rquad = fstep + S.nfracsteps;         // order 2 only

  // ---- rotator tables: FP64 phase, FP32 residual (mode.cpp:909-978, 1054-1097)
  rotator_tables(B, S, G, L, L / 2, tid, NT, true, true, rsub, rstep, fsub, fstep, rquad);
  ts.store(twlo, twhi, tw256, lut8s, tid);
  if (fast) {
#pragma unroll
    for (int q = 0; q < NRAW; q++) {
      const int i = tid + q * NT;
      if (i < RAWW) raw[i] = rw[q];
    }
  }
  if (NBITS != 8 && tid < 4) lut8s[tid] = A.st[B.station].lut[tid];   // (S.lut[tid] would put S in scratch)
  __syncthreads();

  const float4 lut2 = make_float4(S.lut[0], S.lut[1], S.lut[2], S.lut[3]);
  const int s = S.s, slog = S.slog;
  const int bitoff = (int)(bit0 & 31);
  // Strides between a thread's points (j + r L/16) in the padded LDS layout
  // and in the staged stream words; both constant when L/16 is a multiple
  // of 16 (points) and of 32 bits (samples).
  constexpr bool kStride16 = NTF % 16 == 0;
  constexpr int RSTR = NTF + NTF / 16;
  constexpr bool kWordStep = (NBITS * NTF) % 32 == 0;
  constexpr int WSTEP = (NBITS * NTF) / 32;
  // n & (s-1) is the same for all 16 points of a thread: one sub-rotator
  // value per thread, a step-rotator value per point
  const bool rotfast = slog >= 0 && (NTF & (s - 1)) == 0;

  // ---- pass 1 (Ns = 1): decode, fringe-rotate, radix-16 DFT, to LDS
  float2 v[16];
  if (tid < NTF) {
    const int j = tid;
    bool transformed = false;                  // the branch ran the 16-point DFT itself
    if (kWordStep && fast && rotfast && S.order == 1) {
      const int bp = bitoff + NBITS * j;
      const unsigned* wp = raw + (bp >> 5);
      const int sh = bp & 31;
      const float2 r0 = rsub[j & (s - 1)];
      const float2* stp = rstep + (j >> slog);
      const int sstep = NTF >> slog;
      // x[r] rot[r] and x[r + 8] rot[r + 8] go straight into the first
      // butterflies: sum and difference as two FMAs on the first product
      float2 sm[8], df[8];
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const unsigned c0 = (wp[r * WSTEP] >> sh) & ((1u << NBITS) - 1u);
        const unsigned c8 = (wp[(r + 8) * WSTEP] >> sh) & ((1u << NBITS) - 1u);
        const float x0 = lut8s[c0], x8 = lut8s[c8];
        const float2 q0 = cmulf(r0, stp[r * sstep]);         // sub[k] * step[j], mode.cpp:975-977
        const float2 q8 = cmulf(r0, stp[(r + 8) * sstep]);
        const float2 p = make_float2(x0 * q0.x, x0 * q0.y);
        sm[r] = make_float2(__builtin_fmaf(x8, q8.x, p.x), __builtin_fmaf(x8, q8.y, p.y));
        df[r] = make_float2(__builtin_fmaf(-x8, q8.x, p.x), __builtin_fmaf(-x8, q8.y, p.y));
      }
      dft16_from_halves(sm, df, v);
      transformed = true;
    } else if (kWordStep && fast && S.order == 0) {
      // post-F fringe rotation: the samples are transformed unrotated
      const int bp = bitoff + NBITS * j;
      const unsigned* wp = raw + (bp >> 5);
      const int sh = bp & 31;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const unsigned code = (wp[r * WSTEP] >> sh) & ((1u << NBITS) - 1u);
        v[r] = make_float2(lut8s[code], 0.0f);
      }
    } else if (kWordStep && fast && rotfast && S.order == 2) {
      // piecewise quadratic rotator: rot[j s + k] = piece_j[k] step[j] with
      // piece_0 = sub[k], piece_{j+1} = piece_j quad[k] (mode.cpp:1043-1049);
      // this thread's points all have the same k and step index q0 + r sstep
      const int bp = bitoff + NBITS * j;
      const unsigned* wp = raw + (bp >> 5);
      const int sh = bp & 31;
      const int k = j & (s - 1), q0 = j >> slog, sstep = NTF >> slog;
      const float2 qd = rquad[k];
      float2 piece = rsub[k];
      for (int i = 0; i < q0; i++) piece = cmulf(piece, qd);
      const float2* stp = rstep + q0;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const unsigned code = (wp[r * WSTEP] >> sh) & ((1u << NBITS) - 1u);
        const float x = lut8s[code];
        const float2 rot = cmulf(piece, *stp);
        stp += sstep;
        v[r] = make_float2(x * rot.x, x * rot.y);
        if (r < 15)
          for (int i = 0; i < sstep; i++) piece = cmulf(piece, qd);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int n = j + r * NTF;
        float x;
        if (fast) {
          const int bp = bitoff + n * NBITS;
          const unsigned code = (raw[bp >> 5] >> (bp & 31)) & ((1u << NBITS) - 1u);
          x = lut8s[code];
        } else {
          x = sample_value(S, Q, A.ftab, NBITS == 8 ? lut8s : A.lut8, lut2, t0 + n, G.validend, B.band);
        }
        float2 rot;
        if (S.order == 0) {
          rot = make_float2(1.0f, 0.0f);
        } else if (S.order == 1) {
          if (slog >= 0) rot = n < s ? rsub[n] : cmulf(rsub[n & (s - 1)], rstep[n >> slog]);
          else rot = n < s ? rsub[n] : cmulf(rsub[n % s], rstep[n / s]);
        } else {
          const int kk = slog >= 0 ? (n & (s - 1)) : n % s;
          const int jj = slog >= 0 ? (n >> slog) : n / s;
          const float2 qd = rquad[kk];
          float2 piece = rsub[kk];
          for (int i = 0; i < jj; i++) piece = cmulf(piece, qd);
          rot = cmulf(piece, rstep[jj]);
        }
        v[r] = S.order == 0 ? make_float2(x, 0.0f) : make_float2(x * rot.x, x * rot.y);
      }
    }
    if (!transformed) DFT<16>::run(v);
  }
  if (kAliasRaw) __syncthreads();            // every stream word read before buf is written
  if (tid < NTF) {
    float2* o = buf + 17 * tid;              // pidx(16 j + r) = 17 j + r
#pragma unroll
    for (int r = 0; r < 16; r++) o[r] = v[r];
  }
  __syncthreads();

  // ---- middle radix-16 passes: LDS -> registers -> LDS
#pragma unroll
  for (int p = 0; p < Sh::NMID; p++) {
    const int Ns = 16 << (4 * p);
    float2 v[16];
    const int j = tid;
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      float2 w[16];
      if (p == 0) {
        // second pass (Ns = 16): W_L^(r k L/256) = W_256^(r k), r k < 256, read whole
#pragma unroll
        for (int r = 0; r < 16; r++) w[r] = tw256[r * k];
      } else {
        twiddles16(twlo, twhi, k * (L / (Ns * 16)), w);
      }
      const float2* ip = buf + pidx(j);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float2 x = kStride16 ? ip[r * RSTR] : buf[pidx(j + r * NTF)];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      DFT<16>::run(v);
    }
    __syncthreads();
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      const int idxD = (j / Ns) * Ns * 16 + k;
      float2* op = buf + pidx(idxD);           // Ns is a multiple of 16
#pragma unroll
      for (int r = 0; r < 16; r++) op[r * (Ns + Ns / 16)] = v[r];
    }
    __syncthreads();
  }

  // ---- last pass (radix RL, Ns = L/RL): only the bins of the kept sideband
  //      (USB [0, C), real LSB [C, 2C)), then the fractional-sample rotation
  if (RL == 2 && tid < NTF && rotfast) {
    // the sideband is a branch around the whole pass, not a select per output
#define FX_LAST_R2(LSB)                                                                                           \
    if (A.prec == FXG_XMAC_FP32)                                                                                  \
      last_pass_r2<L, LSB, float2*>(buf, twlo, twhi, fsub, fstep, s, slog, 1.0f, (float2*)A.spectra + outbase,    \
                                    A.gstride, tid);                                                              \
    else if (A.prec == FXG_XMAC_FP16)                                                                             \
      last_pass_r2<L, LSB, __half2*>(buf, twlo, twhi, fsub, fstep, s, slog, A.scale,                              \
                                     (__half2*)A.spectra + outbase, A.gstride, tid);                              \
    else if (A.prec == FXG_XMAC_FP8)                                                                              \
      last_pass_r2<L, LSB, Fp8Out>(buf, twlo, twhi, fsub, fstep, s, slog, 1.0f,                                   \
                                   Fp8Out{(unsigned short*)A.spectra + outbase, A.cexp + (long long)gband * C},   \
                                   A.gstride, tid);                                                               \
    else                                                                                                          \
      last_pass_r2<L, LSB, __hip_bfloat162*>(buf, twlo, twhi, fsub, fstep, s, slog, A.scale,                      \
                                             (__hip_bfloat162*)A.spectra + outbase, A.gstride, tid);
    if (B.lsb) { FX_LAST_R2(true) } else { FX_LAST_R2(false) }
#undef FX_LAST_R2
  } else if (tid < NTF) {
    const int j = tid;
    const int mh = B.lsb ? RL / 2 : 0;
#pragma unroll
    for (int u = 0; u < 16 / RL; u++) {
      const int k = j + u * NTF;             // butterfly, k < L/RL
      float2 v[RL], w[RL];
      if (RL == 16) twiddles16(twlo, twhi, k, w);
      else twiddlesR<RL>(twlo, twhi, k, w);
#pragma unroll
      for (int r = 0; r < RL; r++) {
        const float2 x = buf[pidx(k + NSL * r)];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      if (RL == 2) {
        v[0] = mh ? sub2(v[0], v[1]) : add2(v[0], v[1]);
      } else {
        DFT<RL>::run(v);
      }
#pragma unroll
      for (int q = 0; q < (RL == 2 ? 1 : RL / 2); q++) {
        const int c = k + NSL * q;             // output channel
        const float2 x = RL == 2 ? v[0] : v[mh + q];
        const float2 fr = slog >= 0 ? cmulf(fsub[c & (s - 1)], fstep[c >> slog]) : cmulf(fsub[c % s], fstep[c / s]);
        const float2 y = make_float2(x.x * fr.x - x.y * fr.y, x.x * fr.y + x.y * fr.x);
        if (A.prec == FXG_XMAC_FP32) {
          ((float2*)A.spectra)[outbase + cidx(c, A.gstride)] = y;
        } else if (A.prec == FXG_XMAC_FP16) {
          ((__half2*)A.spectra)[outbase + cidx(c, A.gstride)] = __floats2half2_rn(y.x * A.scale, y.y * A.scale);
        } else if (A.prec == FXG_XMAC_FP8) {
          ((unsigned short*)A.spectra)[outbase + cidx(c, A.gstride)] = fp8_pair(y, A.cexp[(long long)gband * C + c]);
        } else {
          ((__hip_bfloat162*)A.spectra)[outbase + cidx(c, A.gstride)] =
              __float22bfloat162_rn(make_float2(y.x * A.scale, y.y * A.scale));
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_fstage_c<L, NBITS>: the F stage of complex-sampled data (Mode with
// usecomplex, mode.cpp:1184-1241): L = nchan complex samples per block, each
// a unit of 2 NBITS bits (real part low), times the fringe rotator, an
// L-point transform with every output kept, channel c of the band taken from
// bin m (USB: m = c; DSB: the halves swapped, m = c + L/2 mod L; LSB:
// conjugated in reverse order, m = L - c with DC kept unconjugated; LSB DSB:
// m = L/2 - c mod L, conjugated), then the fractional-sample rotation.
// ---------------------------------------------------------------------------
template <int L, int NBITS>
__global__ __launch_bounds__(FftShape<L>::NT) void k_fstage_c(FStageArgs A) {
  using Sh = FftShape<L>;
  constexpr int NT = Sh::NT, NTF = Sh::NTF, RL = Sh::RL, NSL = Sh::NSL;
  constexpr int UB = 2 * NBITS;                 // bits per complex sample
  constexpr int RAWW = (L * UB) / 32 + 2;
  constexpr bool kAliasRaw = NBITS == 8 && L >= 16384;
  // one LDS block with the small tables below the FFT buffer (immediate
  // offsets reach them)
  struct Smem {
    float2 tabs[Sh::TABS];
    float2 twlo[64], twhi[L / 64];
    float2 tw256[Sh::NMID > 0 ? 256 : 1];       // e^{-2 pi i m/256}: the second pass's twiddles
    float lut8s[NBITS == 8 ? 256 : 4];
    unsigned raw_own[kAliasRaw ? 2 : RAWW + (RAWW & 1)];
    float2 buf[Sh::PAD];
  };
  __shared__ Smem sm;
  float2 (&buf)[Sh::PAD] = sm.buf;
  float2 (&tabs)[Sh::TABS] = sm.tabs;
  float2 (&twlo)[64] = sm.twlo;
  float2 (&twhi)[L / 64] = sm.twhi;
  float2* tw256 = sm.tw256;
  float (&lut8s)[NBITS == 8 ? 256 : 4] = sm.lut8s;
  unsigned* raw = kAliasRaw ? (unsigned*)(buf + Sh::PAD) - RAWW : sm.raw_own;

  const int tid = threadIdx.x;
  const int cb = (gridDim.x % 8 == 0) ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
  const int blk = A.chunk0 + cb;
  TabStage<L, NT, NBITS, (Sh::NMID > 0)> ts;
  ts.load(A.tw, A.lut8, tid);
  const int gband = A.bandlist ? A.bandlist[blockIdx.y] : blockIdx.y;
  const DevBand B = A.bands[gband];
  const DevStation S = A.st[B.station];
  const DevStationSub Q = A.ss[B.station];
  const DevBlock G = A.blocks[(long long)B.station * A.nblocks_subint + blk];
  constexpr int C = L;
  const long long outbase = B.specoff + (long long)cb * A.bstride;

  if (G.status != 0) {
    zero_spec(A.spectra, A.prec, outbase, C, A.gstride, tid, NT);
    return;
  }
  const long long t0 = G.nearestsample;
  const bool fast = G.allvalid != 0;
  const long long bit0 = t0 * UB;
  // one band of 2-bit complex samples (granularity 2): the unpacker's erasure
  // by float index zeroes window sample L/2 (even start) or the imaginary
  // part of window sample L/2 - 1 (odd start); k_prologue sends such blocks
  // down the general decode path below
  const int nerase = S.granularity > 1 ? L / 2 - (int)(t0 & 1) : -1;
  const bool erase_re = (t0 & 1) == 0;
  constexpr int NRAW = (RAWW + NT - 1) / NT;
  unsigned rw[NRAW];
  if (fast) {                                   // loads now, LDS stores after the rotator tables
    const long long w0 = bit0 >> 5;
    const unsigned* src = (const unsigned*)(A.ct + Q.ctbase + (long long)B.band * Q.ctstride);
    const long long nwords = Q.ctstride >> 2;
#pragma unroll
    for (int q = 0; q < NRAW; q++) {
      const int i = tid + q * NT;
      rw[q] = (i < RAWW && w0 + i < nwords) ? src[w0 + i] : 0u;
    }
  }
  float2* rsub = tabs;
  float2* rstep = tabs + S.s;
  float2* fsub = rstep + S.nfrsteps;
  float2* fstep = fsub + S.s;
  float2* rquad = fstep + S.nfracsteps;
  rotator_tables(B, S, G, L, L, tid, NT, true, true, rsub, rstep, fsub, fstep, rquad);
  ts.store(twlo, twhi, tw256, lut8s, tid);
  if (fast) {
#pragma unroll
    for (int q = 0; q < NRAW; q++) {
      const int i = tid + q * NT;
      if (i < RAWW) raw[i] = rw[q];
    }
  }
  if (NBITS != 8 && tid < 4) lut8s[tid] = A.st[B.station].lut[tid];
  __syncthreads();

  const int s = S.s, slog = S.slog;
  const int bitoff = (int)(bit0 & 31);
  constexpr bool kStride16 = NTF % 16 == 0;
  constexpr int RSTR = NTF + NTF / 16;
  constexpr int WSTEP = (UB * NTF) / 32;        // UB * NTF is a multiple of 32 for L >= 128
  const bool rotfast = slog >= 0 && (NTF & (s - 1)) == 0;

  // ---- pass 1: decode, rotator times sample (vectorMul_cf32), radix 16
  float2 v[16];
  if (tid < NTF) {
    const int j = tid;
    if (fast && rotfast) {
      const int bp = bitoff + UB * j;
      const unsigned* wp = raw + (bp >> 5);
      const int sh = bp & 31;                   // a unit never straddles a word
      const int k = j & (s - 1), q0 = j >> slog, sstep = NTF >> slog;
      float2 piece = rsub[k];
      float2 qd = make_float2(1.0f, 0.0f);
      if (S.order == 2) {
        qd = rquad[k];
        for (int i = 0; i < q0; i++) piece = cmulf(piece, qd);
      }
      const float2* stp = rstep + q0;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const unsigned code = (wp[r * WSTEP] >> sh) & ((1u << UB) - 1u);
        const float2 x = make_float2(lut8s[code & ((1u << NBITS) - 1u)], lut8s[code >> NBITS]);
        const float2 rot = cmulf(piece, *stp);  // sub[k] step[j] (order 2: piece_j[k] step[j])
        stp += sstep;
        v[r] = cmulf(rot, x);
        if (S.order == 2 && r < 15)
          for (int i = 0; i < sstep; i++) piece = cmulf(piece, qd);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int n = j + r * NTF;
        float2 x;
        if (fast) {
          const int bp = bitoff + n * UB;
          const unsigned code = (raw[bp >> 5] >> (bp & 31)) & ((1u << UB) - 1u);
          x = make_float2(lut8s[code & ((1u << NBITS) - 1u)], lut8s[code >> NBITS]);
        } else {
          x = sample_value_c(S, Q, A.ftab, lut8s, t0 + n, G.validend, B.band);
          if (n == nerase) {
            x.y = 0.0f;
            if (erase_re) x.x = 0.0f;
          }
        }
        const int kk = slog >= 0 ? (n & (s - 1)) : n % s;
        const int jj = slog >= 0 ? (n >> slog) : n / s;
        float2 rot;
        if (S.order == 1) {
          rot = n < s ? rsub[n] : cmulf(rsub[kk], rstep[jj]);
        } else {
          const float2 qd = rquad[kk];
          float2 piece = rsub[kk];
          for (int i = 0; i < jj; i++) piece = cmulf(piece, qd);
          rot = cmulf(piece, rstep[jj]);
        }
        v[r] = cmulf(rot, x);
      }
    }
    DFT<16>::run(v);
  }
  if (kAliasRaw) __syncthreads();
  if (tid < NTF) {
    float2* o = buf + 17 * tid;
#pragma unroll
    for (int r = 0; r < 16; r++) o[r] = v[r];
  }
  __syncthreads();

  // ---- middle radix-16 passes
#pragma unroll
  for (int p = 0; p < Sh::NMID; p++) {
    const int Ns = 16 << (4 * p);
    float2 v[16];
    const int j = tid;
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      float2 w[16];
      if (p == 0) {
        // second pass (Ns = 16): W_L^(r k L/256) = W_256^(r k), r k < 256, read whole
#pragma unroll
        for (int r = 0; r < 16; r++) w[r] = tw256[r * k];
      } else {
        twiddles16(twlo, twhi, k * (L / (Ns * 16)), w);
      }
      const float2* ip = buf + pidx(j);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float2 x = kStride16 ? ip[r * RSTR] : buf[pidx(j + r * NTF)];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      DFT<16>::run(v);
    }
    __syncthreads();
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      const int idxD = (j / Ns) * Ns * 16 + k;
      float2* op = buf + pidx(idxD);
#pragma unroll
      for (int r = 0; r < 16; r++) op[r * (Ns + Ns / 16)] = v[r];
    }
    __syncthreads();
  }

  // ---- last pass (radix RL): every output, to its channel
  if (tid < NTF) {
    const int j = tid;
#pragma unroll
    for (int u = 0; u < 16 / RL; u++) {
      const int k = j + u * NTF;
      float2 v[RL], w[RL];
      if (RL == 16) twiddles16(twlo, twhi, k, w);
      else twiddlesR<RL>(twlo, twhi, k, w);
#pragma unroll
      for (int r = 0; r < RL; r++) {
        const float2 x = buf[pidx(k + NSL * r)];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      DFT<RL>::run(v);
#pragma unroll
      for (int q = 0; q < RL; q++) {
        const int m = k + NSL * q;               // transform bin
        int c = m;
        bool cj = false;
        if (!B.lsb) {
          if (B.dsb) c = (m + C / 2) & (C - 1);
        } else if (B.dsb) {
          c = (C / 2 - m) & (C - 1);
          cj = true;
        } else if (m > 0) {
          c = C - m;
          cj = true;
        }
        const float2 x = cj ? make_float2(v[q].x, -v[q].y) : v[q];
        const float2 fr = slog >= 0 ? cmulf(fsub[c & (s - 1)], fstep[c >> slog]) : cmulf(fsub[c % s], fstep[c / s]);
        const float2 y = make_float2(x.x * fr.x - x.y * fr.y, x.x * fr.y + x.y * fr.x);
        if (A.prec == FXG_XMAC_FP32) {
          ((float2*)A.spectra)[outbase + cidx(c, A.gstride)] = y;
        } else if (A.prec == FXG_XMAC_FP16) {
          ((__half2*)A.spectra)[outbase + cidx(c, A.gstride)] = __floats2half2_rn(y.x * A.scale, y.y * A.scale);
        } else if (A.prec == FXG_XMAC_FP8) {
          ((unsigned short*)A.spectra)[outbase + cidx(c, A.gstride)] = fp8_pair(y, A.cexp[(long long)gband * C + c]);
        } else {
          ((__hip_bfloat162*)A.spectra)[outbase + cidx(c, A.gstride)] = __float22bfloat162_rn(make_float2(y.x * A.scale, y.y * A.scale));
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FFT lengths above 16384: L = 16 M and
//     X[k2 + M k1] = sum_n1 W_16^(n1 k1) W_L^(n1 k2) Y_n1[k2],
// Y_n1 the M-point FFT of x[n1 + 16 n2], n2 < M.  k_bigfft_a: one workgroup
// per (block, band, n1) decodes and fringe-rotates its decimated samples,
// transforms them in LDS and writes W_L^(n1 k2) Y_n1[k2] to the scratch
// Z[band][block][n1][M]; k_bigfft_b: one thread per k2 does the 16-point DFT
// over n1, keeps the sideband's eight outputs and applies the
// fractional-sample rotation.
// ---------------------------------------------------------------------------
// rotator tables of each (band, block) of a sub-chunk, once, for the 16
// workgroups of k_bigfft_a and the rows of k_bigfft_b that read them:
// [sub s][step nfrsteps][quad s][frac sub s][frac step nfracsteps]
template <int L>
__global__ __launch_bounds__(256) void k_rotab(FStageArgs A) {
  const int cbl = blockIdx.x;
  const int blk = A.chunk0 + A.zcb0 + cbl;
  const int gband = A.bandlist ? A.bandlist[blockIdx.y] : blockIdx.y;
  const DevBand B = A.bands[gband];
  const DevStation S = A.st[B.station];
  const DevBlock G = A.blocks[(long long)B.station * A.nblocks_subint + blk];
  if (G.status != 0) return;
  float2* rt = A.rtab + ((long long)gband * A.zcap + cbl) * A.rtabn;
  rotator_tables(B, S, G, L, L / 2, threadIdx.x, 256, true, true, rt, rt + S.s, rt + 2 * S.s + S.nfrsteps,
                 rt + 3 * S.s + S.nfrsteps, rt + S.s + S.nfrsteps);
}

template <int M, int NBITS>
__global__ __launch_bounds__(FftShape<M>::NT) void k_bigfft_a(FStageArgs A) {
  using Sh = FftShape<M>;
  constexpr int L = 16 * M;
  constexpr int NT = Sh::NT, NTF = Sh::NTF, RL = Sh::RL, NSL = Sh::NSL;
  static_assert(NTF % 16 == 0, "M/16 threads, a multiple of 16");
  __shared__ float2 buf[Sh::PAD];
  __shared__ float2 tabs[fxg_rotator_table_entries(L)];
  __shared__ float2 twlo[64], twhi[L / 64];
  __shared__ float lut8s[NBITS == 8 ? 256 : 4];

  const int tid = threadIdx.x;
  const int cbl = blockIdx.x, n1 = blockIdx.z;   // block within the sub-chunk, decimation phase
  const int blk = A.chunk0 + A.zcb0 + cbl;
  TabStage<L, NT, NBITS, false> ts;             // loads in flight while the records arrive
  ts.load(A.tw, A.lut8, tid);
  const int gband = A.bandlist ? A.bandlist[blockIdx.y] : blockIdx.y;
  const DevBand B = A.bands[gband];
  const DevStation S = A.st[B.station];
  const DevStationSub Q = A.ss[B.station];
  const DevBlock G = A.blocks[(long long)B.station * A.nblocks_subint + blk];
  if (G.status != 0) return;                     // k_bigfft_b writes the zero spectrum

  // the pre-F rotator tables k_rotab computed for this (band, block)
  float2* rsub = tabs;
  float2* rstep = tabs + S.s;
  float2* rquad = rstep + S.nfrsteps;
  {
    // the rotator table entries, all loads in flight before the LDS stores
    constexpr int NROT = (fxg_rotator_table_entries(L) + NT - 1) / NT;
    const float2* rt = A.rtab + ((long long)gband * A.zcap + cbl) * A.rtabn;
    const int nrot = S.order == 0 ? 0 : (S.order == 2 ? 2 * S.s : S.s) + S.nfrsteps;
    float2 rr[NROT];
#pragma unroll
    for (int q = 0; q < NROT; q++) {
      const int i = tid + q * NT;
      if (i < nrot) rr[q] = rt[i];
    }
    ts.store(twlo, twhi, nullptr, lut8s, tid);
#pragma unroll
    for (int q = 0; q < NROT; q++) {
      const int i = tid + q * NT;
      if (i < nrot) tabs[i] = rr[q];
    }
  }
  if (NBITS != 8 && tid < 4) lut8s[tid] = A.st[B.station].lut[tid];
  __syncthreads();

  const float4 lut2 = make_float4(S.lut[0], S.lut[1], S.lut[2], S.lut[3]);
  const int s = S.s, slog = S.slog;
  const long long t0 = G.nearestsample;
  const bool fast = G.allvalid != 0;
  const unsigned* stream = (const unsigned*)(A.ct + Q.ctbase + (long long)B.band * Q.ctstride);
  const long long nwords = Q.ctstride >> 2;
  const long long bit0 = t0 * NBITS;

  // ---- pass 1 (Ns = 1) over n2 = j + r M/16: decode, rotate, radix-16 DFT
  if (tid < NTF) {
    const int j = tid;
    float2 v[16];
    // samples n1 + 16 (j + r M/16) lie WSTEP whole stream words apart at one
    // bit offset, and (16 M/16 being a multiple of the stride s) share one
    // sub-rotator index, their step indices sstep apart
    constexpr int WSTEP = (16 * NBITS * NTF) / 32;
    const int n0 = n1 + 16 * j;
    const long long bq = bit0 + (long long)NBITS * n0;
    const long long wq = bq >> 5;
    const int sh = (int)(bq & 31);
    const bool wordfast = fast && (S.order == 1 || S.order == 2) && slog >= 0 && ((16 * NTF) & (s - 1)) == 0 &&
                          wq + 15LL * WSTEP < nwords;
    if (wordfast && S.order == 1) {
      const float2 r0 = rsub[n0 & (s - 1)];
      const float2* stp = rstep + (n0 >> slog);
      const int sstep = (16 * NTF) >> slog;
      const unsigned* wp = stream + wq;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float x = lut8s[(wp[r * WSTEP] >> sh) & ((1u << NBITS) - 1u)];
        const float2 rot = cmulf(r0, *stp);     // sub[k] * step[j], mode.cpp:975-977
        stp += sstep;
        v[r] = make_float2(x * rot.x, x * rot.y);
      }
    } else if (wordfast) {
      // piecewise quadratic rotator: piece_q[k] = sub[k] quad[k]^q, one
      // product chain per thread carried from point to point (the same
      // products, in the same order, as rebuilding it for each point)
      const int k = n0 & (s - 1), q0 = n0 >> slog, sstep = (16 * NTF) >> slog;
      const float2 qd = rquad[k];
      float2 piece = rsub[k];
      for (int i = 0; i < q0; i++) piece = cmulf(piece, qd);
      const float2* stp = rstep + q0;
      const unsigned* wp = stream + wq;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float x = lut8s[(wp[r * WSTEP] >> sh) & ((1u << NBITS) - 1u)];
        const float2 rot = cmulf(piece, *stp);
        stp += sstep;
        v[r] = make_float2(x * rot.x, x * rot.y);
        if (r < 15)
          for (int i = 0; i < sstep; i++) piece = cmulf(piece, qd);
      }
    } else {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int n = n1 + 16 * (j + r * NTF);       // sample of the block
      float x;
      if (fast) {
        const long long bp = bit0 + (long long)NBITS * n;
        const unsigned wv = (bp >> 5) < nwords ? stream[bp >> 5] : 0u;
        x = lut8s[(wv >> (bp & 31)) & ((1u << NBITS) - 1u)];
      } else {
        x = sample_value(S, Q, A.ftab, NBITS == 8 ? lut8s : A.lut8, lut2, t0 + n, G.validend, B.band);
      }
      float2 rot = make_float2(1.0f, 0.0f);
      if (S.order != 0) {
        const int kk = slog >= 0 ? (n & (s - 1)) : n % s;
        const int jj = slog >= 0 ? (n >> slog) : n / s;
        if (S.order == 1) {
          rot = n < s ? rsub[n] : cmulf(rsub[kk], rstep[jj]);
        } else {
          const float2 qd = rquad[kk];
          float2 piece = rsub[kk];
          for (int i = 0; i < jj; i++) piece = cmulf(piece, qd);
          rot = cmulf(piece, rstep[jj]);
        }
      }
      v[r] = S.order == 0 ? make_float2(x, 0.0f) : make_float2(x * rot.x, x * rot.y);
    }
    }
    DFT<16>::run(v);
    float2* o = buf + 17 * j;
#pragma unroll
    for (int r = 0; r < 16; r++) o[r] = v[r];
  }
  __syncthreads();

  // ---- middle radix-16 passes (W_M^m = W_L^(16 m))
#pragma unroll
  for (int p = 0; p < Sh::NMID; p++) {
    const int Ns = 16 << (4 * p);
    float2 v[16];
    const int j = tid;
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      float2 w[16];
      twiddles16(twlo, twhi, 16 * k * (M / (Ns * 16)), w);
      constexpr int RSTR = NTF + NTF / 16;     // pidx(j + r NTF) = pidx(j) + r RSTR (NTF a multiple of 16)
      const float2* ip = buf + pidx(j);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float2 x = ip[r * RSTR];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      DFT<16>::run(v);
    }
    __syncthreads();
    if (tid < NTF) {
      const int k = j & (Ns - 1);
      const int idxD = (j / Ns) * Ns * 16 + k;
#pragma unroll
      for (int r = 0; r < 16; r++) buf[pidx(idxD + r * Ns)] = v[r];
    }
    __syncthreads();
  }

  // ---- last pass, every output, times W_L^(n1 k2), to the scratch
  if (tid < NTF) {
    const int j = tid;
    float2* z = A.zbuf + (((long long)gband * A.zcap + cbl) * 16 + n1) * M;
#pragma unroll
    for (int u = 0; u < 16 / RL; u++) {
      const int k = j + u * NTF;
      float2 v[RL], w[RL];
      if (RL == 16) twiddles16(twlo, twhi, 16 * k, w);
      else twiddlesR<RL>(twlo, twhi, 16 * k, w);
#pragma unroll
      for (int r = 0; r < RL; r++) {
        const float2 x = buf[pidx(k + NSL * r)];
        v[r] = r == 0 ? x : cmulf(x, w[r]);
      }
      if (RL == 2) dft2(v[0], v[1]);
      else DFT<RL>::run(v);
#pragma unroll
      for (int m = 0; m < RL; m++) {
        const int k2 = k + NSL * m;
        z[k2] = n1 == 0 ? v[m] : cmulf(v[m], tw_of(twlo, twhi, n1 * k2));
      }
    }
  }
}

template <int M>
__global__ __launch_bounds__(256) void k_bigfft_b(FStageArgs A) {
  constexpr int L = 16 * M, C = L / 2;
  __shared__ float2 tabs[fxg_rotator_table_entries(L)];
  const int tid = threadIdx.x;
  // block within the sub-chunk, consecutive blocks on one XCD (as k_fstage)
  const int cbl = (gridDim.x % 8 == 0) ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
  const int cb = A.zcb0 + cbl;                   // block within the chunk
  const int blk = A.chunk0 + cb;
  const int gband = A.bandlist ? A.bandlist[blockIdx.y] : blockIdx.y;
  const DevBand B = A.bands[gband];
  const DevStation S = A.st[B.station];
  const DevBlock G = A.blocks[(long long)B.station * A.nblocks_subint + blk];
  const int k2 = blockIdx.z * 256 + tid;
  const long long outbase = B.specoff + (long long)cb * A.bstride;
  if (G.status != 0) {
    for (int q = 0; q < 8; q++) {
      if (A.prec == FXG_XMAC_FP32) ((float2*)A.spectra)[outbase + cidx(k2 + M * q, A.gstride)] = make_float2(0.f, 0.f);
      else if (A.prec == FXG_XMAC_FP8) ((unsigned short*)A.spectra)[outbase + cidx(k2 + M * q, A.gstride)] = 0;
      else ((unsigned*)A.spectra)[outbase + cidx(k2 + M * q, A.gstride)] = 0u;
    }
    return;
  }
  float2* fsub = tabs;
  float2* fstep = tabs + S.s;
  // the sub-FFT outputs and the fractional-sample tables: every load in
  // flight before the first LDS store and the barrier
  const float2* z = A.zbuf + ((long long)gband * A.zcap + cbl) * 16 * M + k2;
  float2 v[16];
#pragma unroll
  for (int n1 = 0; n1 < 16; n1++) v[n1] = z[(long long)n1 * M];
  {
    constexpr int NFR = (fxg_rotator_table_entries(L) + 255) / 256;
    const float2* rt = A.rtab + ((long long)gband * A.zcap + cbl) * A.rtabn + 2 * S.s + S.nfrsteps;
    const int nfr = S.s + S.nfracsteps;
    float2 fr[NFR];
#pragma unroll
    for (int q = 0; q < NFR; q++) {
      const int i = tid + q * 256;
      if (i < nfr) fr[q] = rt[i];
    }
#pragma unroll
    for (int q = 0; q < NFR; q++) {
      const int i = tid + q * 256;
      if (i < nfr) tabs[i] = fr[q];
    }
  }
  __syncthreads();
  const int s = S.s, slog = S.slog;
  DFT<16>::run(v);
  const int mh = B.lsb ? 8 : 0;          // USB keeps k1 < 8, real LSB k1 >= 8
  const float scale = (A.prec == FXG_XMAC_FP32 || A.prec == FXG_XMAC_FP8) ? 1.0f : A.scale;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int c = k2 + M * q;
    const float2 x = v[mh + q];
    const float2 fr = slog >= 0 ? cmulf(fsub[c & (s - 1)], fstep[c >> slog]) : cmulf(fsub[c % s], fstep[c / s]);
    const float2 y = make_float2((x.x * fr.x - x.y * fr.y) * scale, (x.x * fr.y + x.y * fr.x) * scale);
    if (A.prec == FXG_XMAC_FP32) ((float2*)A.spectra)[outbase + cidx(c, A.gstride)] = y;
    else if (A.prec == FXG_XMAC_FP16) ((__half2*)A.spectra)[outbase + cidx(c, A.gstride)] = __floats2half2_rn(y.x, y.y);
    else if (A.prec == FXG_XMAC_FP8)
      ((unsigned short*)A.spectra)[outbase + cidx(c, A.gstride)] = fp8_pair(y, A.cexp[(long long)gband * C + c]);
    else ((__hip_bfloat162*)A.spectra)[outbase + cidx(c, A.gstride)] = __float22bfloat162_rn(y);
  }
}

#define INSTANTIATE(Lv) template __global__ void k_fstage<Lv, 2>(FStageArgs); \
                        template __global__ void k_fstage<Lv, 8>(FStageArgs);
INSTANTIATE(128) INSTANTIATE(256) INSTANTIATE(512) INSTANTIATE(1024)
INSTANTIATE(2048) INSTANTIATE(4096) INSTANTIATE(8192) INSTANTIATE(16384)
template __global__ void k_bigfft_a<2048, 2>(FStageArgs);
template __global__ void k_bigfft_a<2048, 8>(FStageArgs);
template __global__ void k_bigfft_a<4096, 2>(FStageArgs);
template __global__ void k_bigfft_a<4096, 8>(FStageArgs);
template __global__ void k_bigfft_b<2048>(FStageArgs);
template __global__ void k_bigfft_b<4096>(FStageArgs);
template __global__ void k_rotab<32768>(FStageArgs);
template __global__ void k_rotab<65536>(FStageArgs);

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// k_frame_digest: a 64-bit digest of each frame's bytes (FNV-1a over the
// bytes taken four at a time, little-endian), one thread per frame, so that
// a caller of the per-block entry point can tell whether the frames a block
// reads changed since they were prepared.
// ---------------------------------------------------------------------------
__host__ __device__ inline unsigned long long frame_digest_bytes(const unsigned char* p, int n) {
  unsigned long long h = 0xcbf29ce484222325ull;
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    const unsigned w = (unsigned)p[i] | ((unsigned)p[i + 1] << 8) | ((unsigned)p[i + 2] << 16) | ((unsigned)p[i + 3] << 24);
    h = (h ^ w) * 0x100000001b3ull;
  }
  for (; i < n; i++) h = (h ^ p[i]) * 0x100000001b3ull;
  return h;
}

__global__ __launch_bounds__(256) void k_frame_digest(const unsigned char* __restrict__ data, long long f0,
                                                      long long nf, int framebytes,
                                                      unsigned long long* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  out[i] = frame_digest_bytes(data + (f0 + i) * (long long)framebytes, framebytes);
}

hipError_t launch_frame_digest(const unsigned char* data, long long f0, long long nf, int framebytes,
                               unsigned long long* out, hipStream_t stream) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_frame_digest, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, stream, data, f0, nf,
                     framebytes, out);
  return hipGetLastError();
}

unsigned long long host_frame_digest(const unsigned char* p, int n) { return frame_digest_bytes(p, n); }

hipError_t launch_frames(const DevStation* st, const DevStationSub* ss, int2* ftab, int nst,
                         long long maxframes, hipStream_t stream) {
  if (maxframes <= 0) return hipSuccess;
  dim3 grid((unsigned)((maxframes + 255) / 256), nst);
  hipLaunchKernelGGL(k_frames, grid, dim3(256), 0, stream, st, ss, ftab);
  return hipGetLastError();
}

hipError_t launch_prologue(const DevStation* st, const DevStationSub* ss, const int* flags,
                           const int2* ftab, int nblocks, DevBlock* out, int nst, hipStream_t stream) {
  dim3 grid((unsigned)((nblocks + 127) / 128), nst);
  hipLaunchKernelGGL(k_prologue, grid, dim3(128), 0, stream, st, ss, flags, ftab, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_fstage(int L, int nbits, int cplx, const FStageArgs& a, int nbands, int nblocks, hipStream_t stream) {
  dim3 grid(nblocks, nbands);
  if (nbits != 2 && nbits != 8) return hipErrorInvalidValue;
  if (cplx) {
    switch (L) {
#define CASE(Lv) case Lv: \
      if (nbits == 2) hipLaunchKernelGGL((k_fstage_c<Lv, 2>), grid, dim3(FftShape<Lv>::NT), 0, stream, a); \
      else hipLaunchKernelGGL((k_fstage_c<Lv, 8>), grid, dim3(FftShape<Lv>::NT), 0, stream, a); \
      break;
      CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096) CASE(8192) CASE(16384)
#undef CASE
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  switch (L) {
#define CASE(Lv) case Lv: \
      if (nbits == 2) hipLaunchKernelGGL((k_fstage<Lv, 2>), grid, dim3(FftShape<Lv>::NT), 0, stream, a); \
      else hipLaunchKernelGGL((k_fstage<Lv, 8>), grid, dim3(FftShape<Lv>::NT), 0, stream, a); \
      break;
    CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096) CASE(8192) CASE(16384)
#undef CASE
#define BIG(Mv) case 16 * Mv: { \
      if (!a.zbuf || !a.rtab || a.zcap < 1) return hipErrorInvalidValue; \
      for (int zc = 0; zc < nblocks; zc += a.zcap) { \
        FStageArgs b = a; \
        b.zcb0 = zc; \
        const int n = nblocks - zc < a.zcap ? nblocks - zc : a.zcap; \
        dim3 gr(n, nbands), ga(n, nbands, 16), gb(n, nbands, Mv / 256); \
        hipLaunchKernelGGL((k_rotab<16 * Mv>), gr, dim3(256), 0, stream, b); \
        if (nbits == 2) hipLaunchKernelGGL((k_bigfft_a<Mv, 2>), ga, dim3(FftShape<Mv>::NT), 0, stream, b); \
        else hipLaunchKernelGGL((k_bigfft_a<Mv, 8>), ga, dim3(FftShape<Mv>::NT), 0, stream, b); \
        hipLaunchKernelGGL((k_bigfft_b<Mv>), gb, dim3(256), 0, stream, b); \
      } \
      break; }
    BIG(2048) BIG(4096)
#undef BIG
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cornerturn(const DevStation* st, const DevStationSub* ss, unsigned char* ct, int nst,
                             long long maxframes, int maxdatabytes, int fastW, long long maxsamples,
                             hipStream_t stream) {
  if (maxframes <= 0) return hipSuccess;
  if (fastW) {
    const long long groups = (maxsamples + 15) / 16;
    dim3 grid((unsigned)((groups + 255) / 256), nst);
    if (fastW == 1) hipLaunchKernelGGL(k_cornerturn2<1>, grid, dim3(256), 0, stream, st, ss, ct);
    else if (fastW == 2) hipLaunchKernelGGL(k_cornerturn2<2>, grid, dim3(256), 0, stream, st, ss, ct);
    else hipLaunchKernelGGL(k_cornerturn2<4>, grid, dim3(256), 0, stream, st, ss, ct);
    return hipGetLastError();
  }
  dim3 grid((unsigned)maxframes, nst);
  hipLaunchKernelGGL(k_cornerturn, grid, dim3(256), (size_t)maxdatabytes, stream, st, ss, ct);
  return hipGetLastError();
}

hipError_t launch_unpack(const DevStation* st, const DevStationSub* ss, const int2* ftab, const float* lut8,
                         long long first, int nsamp, int nchan, long long validend, float* out,
                         hipStream_t stream) {
  dim3 grid((unsigned)((nsamp + 255) / 256), nchan);
  hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, stream, st, ss, ftab, lut8, first, nsamp, validend, out);
  return hipGetLastError();
}

}  // namespace fxg
