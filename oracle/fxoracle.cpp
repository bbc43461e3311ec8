/*
 * fxoracle.cpp — CPU restatement of the DiFX mpifxcorr FX hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libfxcorr_gpu.so) never links or calls it.
 *
 * It restates, with GENERIC-architecture semantics (architecture.h.in:404-1075:
 * libm sinf/cosf, unnormalised forward FFT, truncating int casts, no FMA
 * contraction), the following reference code:
 *
 *   unpack   libraries/mark5access/mark5access/format_vdif.c:74-149 (LUTs),
 *            :768-1303 (2-bit decoders), :1837-2020 (8-bit decoders),
 *            :3840-3848 (granularity), :4009-4066 (validate);
 *            mark5_format_mark5b.c:59-96 (LUT), :1579-1950 (2-bit decoders);
 *            mark5_stream_unpacker.c:129-148 (unpack_with_offset);
 *            mark5_stream.c:153-202 (next_frame); blanker_mark5.c:53-165,282-312;
 *            mpifxcorr/src/mk5mode.cpp:104-177 (Mk5Mode::unpack)
 *   station  mpifxcorr/src/mode.cpp:43-462 (constructor tables), :649-1397
 *            (Mode::process, fringe orders 0, 1, 2, real sampling)
 *   model    mpifxcorr/src/model.cpp:229-304 (calculateDelayInterpolator)
 *   xmac     mpifxcorr/src/core.cpp:667-1122 (processdata), :1159-1369
 *            (averageAndSendAutocorrs), :1586-1867 (uvshiftAndAverage, single
 *            phase centre, no channel averaging)
 *   layout   mpifxcorr/src/configuration.cpp:46-59, 1771-1845, 2212-2441, 2464-2546
 *   pcal     mpifxcorr/src/mode.cpp:412-428, 774-788 (extraction per block);
 *            pcal.cpp:139-229 (getNew, calcNumTones), :253-331 (the direct
 *            reference method), :333-470 (first tone at DC), :498-690
 *            (shifting), :927-1104 (implicit shift), :707-925 and
 *            :1107-1285 (complex samples: shifting, implicit shift),
 *            :1376-1450 (dummy);
 *            mathutil.cpp:8-30 (gcd); core.cpp:1124-1157 (copyPCalTones)
 *
 * Two arithmetic flavours share one code path through a template parameter:
 *   float  — the reference's own FP32 arithmetic (the drop-in parity target);
 *   double — the same algorithm in FP64 end to end (the error yardstick for
 *            the reduced-precision cross-multiply).
 * The FP64 block geometry (delays, nearest sample, integer delay) is the same
 * FP64 expression sequence in both.
 */
#include "../include/fxcorr_gpu.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <complex>
#include <algorithm>
#include <memory>

namespace orc {

constexpr double TWO_PI = 6.283185307179586476925286766559; /* architecture.h.in:435 */
constexpr double TINY = 0.000000001;                          /* mode.cpp:35 */
constexpr float HIMAG = 3.3359f;                              /* mark5_stream.h:75 */

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) { g_err = msg; return code; }

/* ------------------------------------------------------------------------ */
/* Lookup tables                                                            */
/* ------------------------------------------------------------------------ */
/* VDIF 2-bit: code k -> {-HiMag, -1, +1, +HiMag}[k] (format_vdif.c:78,104-107) */
static void lut_vdif2(float* t) { t[0] = -HIMAG; t[1] = -1.0f; t[2] = 1.0f; t[3] = HIMAG; }
/* VDIF 8-bit: (b-128)/3.3 evaluated in double, stored as float (format_vdif.c:126) */
static void lut_vdif8(float* t) { for (int b = 0; b < 256; b++) t[b] = (b - 128) / 3.3; }
/* Mark5B 2-bit: sign bit at even position, magnitude at odd, table
 * {-HiMag, 1, -1, HiMag}[s + 2m] (mark5_format_mark5b.c:63,80-83) */
static void lut_mark5b2(float* t) { t[0] = -HIMAG; t[1] = 1.0f; t[2] = -1.0f; t[3] = HIMAG; }

/* ------------------------------------------------------------------------ */
/* Stride rules (configuration.cpp:46-59, 2464-2511)                          */
/* ------------------------------------------------------------------------ */
static int calcstridelength(unsigned int arraylength) {
  if (arraylength <= 16) return (int)arraylength;
  for (unsigned int guess = (unsigned int)std::sqrt(arraylength + 0.1);; --guess)
    if (arraylength % guess == 0) return (int)(arraylength / guess);
}
static long gcdl(long a, long b) { while (b) { long t = a % b; a = b; b = t; } return a; }

/* ------------------------------------------------------------------------ */
/* Frame-stream decoder (mark5access restated)                              */
/* ------------------------------------------------------------------------ */
struct StreamFmt {
  int format, nbits, nchan, framebytes, headerbytes, databytes, blanker;
  bool cplx;             /* complex samples: 2*nbits per unit, real part low */
  int unitbits;          /* nbits, or 2*nbits for complex */
  int framesamples;      /* sample-times per frame */
  int granularity;       /* samplegranularity (format_vdif.c:3845; mark5b.c:2837) */
  int groupbytes;        /* bytes consumed per decoder iteration */
  int groupsamples;      /* sample-times produced per decoder iteration */
  bool noadvance_quirk;  /* 2/3/4-channel 8-bit VDIF decoders do not advance on blank */
  bool edv4;             /* INTERLACEDVDIF: per-band weights from EDV4 validity masks */
  float lut[256];
};

static int make_fmt(const fxg_datastream* ds, StreamFmt* f) {
  /* INTERLACEDVDIF decodes as VDIF; Mk5Mode adds the EDV4 per-band blanking
   * and weights (mk5mode.cpp:78-90, 145-162) */
  f->edv4 = ds->format == FXG_FORMAT_INTERLACEDVDIF;
  if (f->edv4 && ds->sampling != FXG_SAMPLING_REAL)
    return fail(FXG_EUNSUPPORTED, "INTERLACEDVDIF with complex sampling");
  f->format = f->edv4 ? FXG_FORMAT_VDIF : ds->format; f->nbits = ds->nbits; f->nchan = ds->nrecbands;
  f->framebytes = ds->framebytes; f->headerbytes = ds->headerbytes;
  f->databytes = ds->framebytes - ds->headerbytes; f->blanker = ds->blanker;
  if (f->nchan <= 0 || (f->nchan & (f->nchan - 1)) != 0)
    return fail(FXG_EUNSUPPORTED, "nrecbands must be a power of two");
  f->cplx = ds->sampling == FXG_SAMPLING_COMPLEX;
  if (ds->sampling != FXG_SAMPLING_REAL && !f->cplx) return fail(FXG_EINVAL, "sampling");
  if (f->cplx && ds->format != FXG_FORMAT_VDIF) return fail(FXG_EUNSUPPORTED, "complex sampling is VDIF only");
  f->unitbits = f->cplx ? 2 * f->nbits : f->nbits;
  int bitsper = f->nchan * f->unitbits;
  /* complex units smaller than a byte (1-channel 2-bit) have a sample
   * granularity of 2, whose erasure in Mk5Mode::unpack (mk5mode.cpp:121-140)
   * runs over the complex array's floats as if they were real samples
   * (mk5_unpack below does the same) */
  if (f->format == FXG_FORMAT_VDIF) {
    if (f->nbits == 2) lut_vdif2(f->lut);
    else if (f->nbits == 8) lut_vdif8(f->lut);
    else return fail(FXG_EUNSUPPORTED, "VDIF nbits must be 2 or 8");
  } else if (ds->format == FXG_FORMAT_MARK5B) {
    if (f->nbits != 2) return fail(FXG_EUNSUPPORTED, "Mark5B nbits must be 2");
    if (f->databytes != 10000 || f->headerbytes != 16)
      return fail(FXG_EINVAL, "Mark5B frames are 16 + 10000 bytes");
    if (bitsper > 32) return fail(FXG_EUNSUPPORTED, "Mark5B bitstreams <= 32");
    lut_mark5b2(f->lut);
  } else return fail(FXG_EINVAL, "unknown format");
  f->granularity = bitsper >= 8 ? 1 : 8 / bitsper;
  f->framesamples = (int)((int64_t)f->databytes * 8 / bitsper);
  f->groupbytes = bitsper >= 8 ? bitsper / 8 : 1;
  f->groupsamples = bitsper >= 8 ? 1 : 8 / bitsper;
  /* decoders that do not advance past a blank group: real 8-bit with 2..4
   * channels (format_vdif.c:1837-2020), complex 8-bit with any channel count
   * (format_vdif.c:2829-3054) */
  f->noadvance_quirk = f->format == FXG_FORMAT_VDIF && f->nbits == 8 &&
                       (f->cplx || (f->nchan >= 2 && f->nchan <= 4));
  return 0;
}

/* little-endian 64-bit fill word 0x1122334411223344 (blanker_mark5.c:42-46) */
static const uint64_t FILL64 = 0x1122334411223344ULL;
static const uint32_t FILL32 = 0x11223344UL;
static uint64_t rd64(const uint8_t* p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
static uint32_t rd32(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }

/* binary searches of blanker_mark5.c:53-123 (word indices relative to payload) */
static int findfirstinvalid(const uint8_t* pl, int start, int end) {
  int n = end - start, i = n / 2, r = n - 1;
  for (int s = n / 4; s > 0; s >>= 1) {
    if (rd64(pl + 8 * (size_t)(start + i)) == FILL64) { r = i; i -= s; } else i += s;
  }
  r = 8 * (start + r);
  if (rd32(pl + 4 * (size_t)(r / 4 - 1)) == FILL32) r -= 4;
  return r;
}
static int findfirstvalid(const uint8_t* pl, int start, int end) {
  int n = end - start, i = n / 2, r = n - 1;
  for (int s = n / 4; s > 0; s >>= 1) {
    if (rd64(pl + 8 * (size_t)(start + i)) != FILL64) { r = i; i -= s; } else i += s;
  }
  r = 8 * (start + r);
  if (rd32(pl + 4 * (size_t)(r / 4 - 2)) == FILL32) r += 4;
  return r;
}

struct FrameState {
  const uint8_t* frame;   /* null when the frame is not inside the buffer */
  int startvalid, endvalid;
};

/* mark5_stream_next_frame (mark5_stream.c:153-202) for the memory unpacker:
 * validate, then blank (whole frame) or run the blanker. */
static void load_frame(const StreamFmt& f, const uint8_t* data, int64_t nbytes,
                       int64_t frameidx, FrameState* st) {
  int64_t start = frameidx * (int64_t)f.framebytes;
  if (frameidx < 0 || start + f.framebytes > nbytes) {
    /* outside the buffer: the reference reads past its buffer here; this
     * restatement defines such frames as invalid */
    st->frame = nullptr; st->startvalid = 1 << 30; st->endvalid = 0; return;
  }
  const uint8_t* fr = data + start;
  st->frame = fr;
  bool valid = true;
  if (f.format == FXG_FORMAT_VDIF) {
    /* mark5_format_vdif_validate (format_vdif.c:4009-4060); the time check is
     * skipped because the unpacker stream has mjd == 0 (mark5_stream.c:301) */
    uint32_t w0 = rd32(fr), w2 = rd32(fr + 8);
    if (w2 == 0) valid = false;
    if ((w0 >> 31) & 1u) valid = false;
  }
  if (!valid) { st->startvalid = 1 << 30; st->endvalid = 0; return; }
  const uint8_t* pl = fr + f.headerbytes;
  int nword = f.databytes / 8;
  if (f.blanker == FXG_BLANKER_NONE) {            /* blanker_none.c:36-43 */
    st->startvalid = 0; st->endvalid = 1 << 30;
  } else if (f.format == FXG_FORMAT_VDIF) {        /* blanker_vdif, blanker_mark5.c:282-312 */
    st->startvalid = 0;
    if (rd64(pl) == FILL64 || rd64(pl + 8 * (size_t)(nword - 1)) == FILL64) st->endvalid = 0;
    else st->endvalid = 1 << 30;
  } else {                                         /* blanker_mark5 zone 0 (one zone: 10000 B < 32 KiB) */
    int b = 0, e = nword;
    bool startOK = rd64(pl + 8 * (size_t)b) != FILL64;
    bool endOK = rd64(pl + 8 * (size_t)(e - 1)) != FILL64;
    if (startOK && endOK) { st->startvalid = 0; st->endvalid = 1 << 30; }
    else if (!startOK && !endOK) { st->startvalid = 1 << 30; st->endvalid = 0; }
    else if (startOK) { st->startvalid = 0; st->endvalid = findfirstinvalid(pl, b, e); }
    else { st->startvalid = findfirstvalid(pl, b, e); st->endvalid = 1 << 30; }
  }
}

/* mark5_unpack_with_offset (mark5_unpack_complex_with_offset) + the format
 * decoder loop.  out[c*nsamp + o] (complex: float pairs at 2*(c*nsamp + o)).
 * Returns goodsamples (the decoder's return value). */
static int decode(const StreamFmt& f, const uint8_t* data, int64_t nbytes,
                  int64_t offsetsamples, int nsamp, float* out) {
  int64_t frameidx = offsetsamples / f.framesamples;
  FrameState st; load_frame(f, data, nbytes, frameidx, &st);
  int i = (int)((offsetsamples % f.framesamples) * f.nchan * f.unitbits / 8);
  int nblank = 0;
  const uint32_t mask = (1u << f.nbits) - 1u;
  const int nc = f.cplx ? 2 : 1;
  for (int o = 0; o < nsamp; o += f.groupsamples) {
    bool blank = (i < st.startvalid) || (i >= st.endvalid) || (st.frame == nullptr);
    if (!blank && f.format == FXG_FORMAT_MARK5B && (st.frame[f.headerbytes - 11] & 0x80)) blank = true;
    if (blank) {
      nblank++;
      for (int k = 0; k < f.groupsamples; k++)
        for (int c = 0; c < f.nchan; c++)
          for (int q = 0; q < nc; q++)
            if (o + k < nsamp) out[((size_t)c * nsamp + o + k) * nc + q] = 0.0f;
      if (!f.noadvance_quirk) i += f.groupbytes;
    } else {
      const uint8_t* g = st.frame + f.headerbytes + i;
      for (int k = 0; k < f.groupsamples; k++)
        for (int c = 0; c < f.nchan; c++)
          for (int q = 0; q < nc; q++) {
            /* component q of the unit: real in the low nbits, imaginary above
             * (complex_lut2bit / lut8bit pairs, format_vdif.c:133-140, 2829-2850) */
            int bit = (k * f.nchan + c) * f.unitbits + q * f.nbits;
            uint32_t code = (g[bit >> 3] >> (bit & 7)) & mask;
            if (o + k < nsamp) out[((size_t)c * nsamp + o + k) * nc + q] = f.lut[code];
          }
      i += f.groupbytes;
    }
    if (i >= f.databytes) { frameidx++; load_frame(f, data, nbytes, frameidx, &st); i = 0; }
  }
  return nsamp - f.groupsamples * nblank;
}

/* blank_vdif_EDV4 (format_vdif.c:4699-4775) over the unpacked window:
 * work[c*stride + i] holds sample unpackstart + i of channel c, i < nsamp.
 * The first frame of the buffer decides whether the stream carries EDV4
 * validity (extended version 4, mask length 1..64); then every frame of the
 * window whose invalid bit is clear and whose validity mask is not complete
 * has the samples of each channel whose mask bit (channel / (nchan / mask
 * length)) is clear set to zero, counting those that were nonzero into
 * invalid[c].  Returns false (invalid[] all zero) when the first frame is
 * not EDV4.  Frames not wholly inside the buffer are skipped (the
 * reference reads their headers past its buffer); masks are 64-bit. */
static bool blank_edv4(const StreamFmt& f, const uint8_t* data, int64_t nbytes, int64_t unpackstart,
                       float* work, int stride, int nsamp, std::vector<int>& invalid) {
  invalid.assign((size_t)f.nchan, 0);
  if (nbytes < 32) return false;
  const uint32_t w4 = rd32(data + 16);
  const int eversion = (int)(w4 >> 24), masklength = (int)((w4 >> 16) & 0xFF);
  if (eversion != 4 || masklength < 1 || masklength > 64) return false;
  const uint32_t w2 = rd32(data + 8), w3 = rd32(data + 12);
  const int nChan = 1 << ((w2 >> 24) & 0x1F);
  const int64_t frameLength = (int64_t)(w2 & 0xFFFFFF) * 8;
  const int nbits = (int)((w3 >> 26) & 0x1F) + 1;
  /* (a first header whose frame length, channels or bits differ from the
   * datastream's is taken as carrying no EDV4 validity here) */
  if (frameLength != f.framebytes || nChan != f.nchan || nbits != f.nbits) return false;
  const int64_t sampPerFrame = (frameLength - 32) * 8 / (nChan * nbits);
  const uint64_t goodMask = nChan >= 64 ? ~0ULL : ((1ULL << nChan) - 1ULL);
  const int d = nChan / masklength;
  if (sampPerFrame <= 0 || d <= 0) return false;
  int64_t frameNum = unpackstart / sampPerFrame;
  int64_t start = unpackstart % sampPerFrame;
  int64_t end = start + nsamp < sampPerFrame ? start + nsamp : sampPerFrame;
  int64_t left = nsamp, v = 0;
  for (;;) {
    const int64_t N = end - start;
    const int64_t off = frameNum * frameLength;
    if (frameNum >= 0 && off + frameLength <= nbytes) {
      const uint8_t* h = data + off;
      const bool hinv = (rd32(h) >> 31) & 1u;
      const uint64_t vmask = rd64(h + 24);
      if (!hinv && vmask != goodMask)
        for (int c = 0; c < nChan && c < f.nchan; ++c)
          if (((vmask >> (c / d)) & 1ULL) == 0)
            for (int64_t q = 0; q < N; ++q) {
              float& x = work[(size_t)c * stride + (size_t)(q + v)];
              if (x != 0.0f) { x = 0.0f; ++invalid[(size_t)c]; }
            }
    }
    left -= N;
    if (left <= 0) break;
    v += N;
    ++frameNum;
    start = 0;
    end = left < sampPerFrame ? left : sampPerFrame;
  }
  return true;
}

/* Mk5Mode::unpack (mk5mode.cpp:104-177).  `work` holds
 * nchan*(unpacksamples+granularity) samples (complex: float pairs); on return
 * the L samples starting at sampleoffset for channel c are at
 * work[c*stride + munged + n] (complex: that sample index times 2).
 * INTERLACEDVDIF: bandweight[c] = (goodsamples - invalid[c]) / L, and the
 * block weight loses the mean invalid count (mk5mode.cpp:145-162). */
struct Unpacked { int stride; int munged; float weight; std::vector<float> bandweight; };
static Unpacked mk5_unpack(const StreamFmt& f, const uint8_t* data, int64_t nbytes,
                           int64_t sampleoffset, int unpacksamples, std::vector<float>& work) {
  int samplestounpack = unpacksamples + (f.granularity > 1 ? f.granularity : 0);
  work.assign((size_t)f.nchan * samplestounpack * (f.cplx ? 2 : 1), 0.0f);
  int64_t unpackstart = sampleoffset - (sampleoffset % f.granularity);
  float goodsamples = (float)decode(f, data, nbytes, unpackstart, samplestounpack, work.data());
  int munged = 0;
  if (f.granularity > 1) {
    /* index i runs over the band array's floats: for complex samples (the
     * array viewed as floats, real then imaginary) the erasure reaches the
     * real part of sample 0 and the floats of sample (unpacksamples +
     * munged) / 2, inside the block's window */
    int erased = 0;
    const size_t bstride = (size_t)samplestounpack * (f.cplx ? 2 : 1);
    munged = (int)(sampleoffset % f.granularity);
    for (int i = 0; i < munged; i++)
      for (int b = 0; b < f.nchan; b++) {
        float& v = work[(size_t)b * bstride + i];
        if (v != 0.0f) { v = 0.0f; erased++; }
      }
    for (int i = unpacksamples + munged; i < samplestounpack; i++)
      for (int b = 0; b < f.nchan; b++) {
        float& v = work[(size_t)b * bstride + i];
        if (v != 0.0f) { v = 0.0f; erased++; }
      }
    goodsamples -= erased / (float)f.nchan;
  }
  Unpacked u; u.stride = samplestounpack; u.munged = munged;
  if (f.edv4 && !f.cplx) {
    std::vector<int> invalid;
    blank_edv4(f, data, nbytes, unpackstart, work.data(), samplestounpack, samplestounpack, invalid);
    int totalinvalid = 0;
    u.bandweight.assign((size_t)f.nchan, 0.0f);
    for (int b = 0; b < f.nchan; ++b) {
      u.bandweight[(size_t)b] = (goodsamples - invalid[(size_t)b]) / (float)unpacksamples;
      totalinvalid += invalid[(size_t)b];
    }
    goodsamples -= (float)totalinvalid / (float)f.nchan;
  }
  if (goodsamples < 0) goodsamples = 0;
  u.weight = goodsamples / (float)unpacksamples;
  return u;
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2, forward, unnormalised, accurate twiddles         */
/* ------------------------------------------------------------------------ */
template <typename T> struct FFT {
  int n = 0, logn = 0;
  std::vector<int> rev;
  std::vector<std::complex<T>> tw;
  explicit FFT(int L) : n(L) {
    while ((1 << logn) < n) logn++;
    rev.resize(n);
    for (int i = 0; i < n; i++) { int r = 0; for (int b = 0; b < logn; b++) if (i & (1 << b)) r |= 1 << (logn - 1 - b); rev[i] = r; }
    tw.resize(n / 2);
    for (int k = 0; k < n / 2; k++) {
      double ang = -TWO_PI * (double)k / (double)n;
      tw[k] = std::complex<T>((T)std::cos(ang), (T)std::sin(ang));
    }
  }
  void run(std::complex<T>* a) const {
    for (int i = 0; i < n; i++) if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len)
        for (int j = 0; j < half; j++) {
          std::complex<T> w = tw[(size_t)j * step];
          std::complex<T> u = a[i + j];
          std::complex<T> x = a[i + j + half];
          T vr = x.real() * w.real() - x.imag() * w.imag();
          T vi = x.real() * w.imag() + x.imag() * w.real();
          a[i + j] = std::complex<T>(u.real() + vr, u.imag() + vi);
          a[i + j + half] = std::complex<T>(u.real() - vr, u.imag() - vi);
        }
    }
  }
};

/* sin/cos in the flavour's precision: libm sinf/cosf for float (GENERIC
 * vectorSinCos_f32, architecture.h.in:~950) and sin/cos for double. */
static inline void sincos_t(float x, float* s, float* c) { *s = sinf(x); *c = cosf(x); }
static inline void sincos_t(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }

/* complex helpers written out so the operation order is that of the generic
 * vector routines (architecture.h.in:580-586, 673-719) */
template <typename T> struct cpx { T re, im; };
template <typename T> static inline cpx<T> cmul(cpx<T> a, cpx<T> v) {
  cpx<T> d; d.re = a.re * v.re - a.im * v.im; d.im = a.re * v.im + a.im * v.re; return d;
}

/* ------------------------------------------------------------------------ */
/* Job geometry                                                              */
/* ------------------------------------------------------------------------ */
struct DsGeom {
  StreamFmt fmt;
  bool cplx, dsb;                 /* complex sampling; double sideband */
  int C, L, s, numfrstrides, nfracsteps;
  double bw, sampletime;          /* MHz, µs */
  int order;                      /* fringe rotation order 0, 1, 2 */
  bool fractionalLoFreq, deltapoloffsets;
  std::vector<double> polre, polim;  /* per recorded band: phasecorrA of the band, or (1, 0) */
  int bytesperblocknum, bytesperblockden, samplesperblock;
};

struct Layout {
  int nfreqs, nds, nbl;
  std::vector<char> freqused, equivused;
  std::vector<std::vector<int64_t>> bloff, bwoff;  /* [freq][baseline], -1 if absent */
  std::vector<int64_t> acoff, acwoff, pcaloff;
  std::vector<std::vector<int>> pcal_ntones, pcal_offhz;  /* [ds][recfreq]; empty: no phase cal */
  std::vector<std::vector<int>> blfreq_local;      /* [baseline][freq] -> local index or -1 */
  int64_t total = 0;
  int bandsperautocorr = 1;
  int xmacstridelen = 0, rotatestridelen = 0;
  std::vector<int> arraystridelen;
};

static int ds_band_freqindex(const fxg_datastream& d, int band) {
  if (band < d.nrecbands) return d.recfreq_index[d.recband_freq[band]];
  return d.zoomband_freqindex[band - d.nrecbands];
}

static int validate(const fxg_job_desc* j) {
  if (!j || !j->freqs || !j->datastreams || (j->nbaselines > 0 && !j->baselines))
    return fail(FXG_EINVAL, "null descriptor");
  if (j->nfreqs <= 0 || j->nfreqs > FXG_MAX_FREQS) return fail(FXG_EINVAL, "nfreqs");
  if (j->ndatastreams <= 0) return fail(FXG_EINVAL, "ndatastreams");
  if (j->blockspersend <= 0) return fail(FXG_EINVAL, "blockspersend");
  if (j->fringe_order < 0 || j->fringe_order > 2) return fail(FXG_EUNSUPPORTED, "fringe rotation order must be 0, 1 or 2");
  for (int f = 0; f < j->nfreqs; f++) {
    if (j->freqs[f].chans_to_avg < 1 || j->freqs[f].nchan % j->freqs[f].chans_to_avg)
      return fail(FXG_EINVAL, "chans_to_avg must divide nchan");
    if (j->freqs[f].nchan <= 0) return fail(FXG_EINVAL, "nchan");
  }
  for (int d = 0; d < j->ndatastreams; d++) {
    const fxg_datastream& ds = j->datastreams[d];
    if (ds.nrecfreqs <= 0 || ds.nrecfreqs > FXG_MAX_RECFREQS) return fail(FXG_EINVAL, "nrecfreqs");
    if (ds.nrecbands <= 0 || ds.nrecbands > FXG_MAX_BANDS) return fail(FXG_EINVAL, "nrecbands");
    if (ds.nzoombands < 0 || ds.nzoombands > FXG_MAX_ZOOMBANDS) return fail(FXG_EINVAL, "nzoombands");
    for (int i = 0; i < ds.nrecfreqs; i++)
      if (ds.recfreq_index[i] < 0 || ds.recfreq_index[i] >= j->nfreqs) return fail(FXG_EINVAL, "recfreq_index");
    for (int b = 0; b < ds.nrecbands; b++)
      if (ds.recband_freq[b] < 0 || ds.recband_freq[b] >= ds.nrecfreqs) return fail(FXG_EINVAL, "recband_freq");
    for (int z = 0; z < ds.nzoombands; z++) {
      if (ds.zoomband_freqindex[z] < 0 || ds.zoomband_freqindex[z] >= j->nfreqs) return fail(FXG_EINVAL, "zoomband_freqindex");
      if (ds.zoomband_parent[z] < 0 || ds.zoomband_parent[z] >= ds.nrecbands) return fail(FXG_EINVAL, "zoomband_parent");
    }
    const fxg_freq& f0 = j->freqs[ds.recfreq_index[0]];
    for (int i = 1; i < ds.nrecfreqs; i++) {
      const fxg_freq& fi = j->freqs[ds.recfreq_index[i]];
      if (fi.nchan != f0.nchan || fi.bw_mhz != f0.bw_mhz) return fail(FXG_EUNSUPPORTED, "recorded freqs of a datastream must share nchan and bandwidth");
    }
  }
  for (int b = 0; b < j->nbaselines; b++) {
    const fxg_baseline& bl = j->baselines[b];
    if (bl.ds1 < 0 || bl.ds1 >= j->ndatastreams || bl.ds2 < 0 || bl.ds2 >= j->ndatastreams) return fail(FXG_EINVAL, "baseline datastream");
    if (bl.nfreqs < 0 || bl.nfreqs > FXG_MAX_BLFREQS) return fail(FXG_EINVAL, "baseline nfreqs");
    for (int f = 0; f < bl.nfreqs; f++) {
      if (bl.freq_index[f] < 0 || bl.freq_index[f] >= j->nfreqs) return fail(FXG_EINVAL, "baseline freq_index");
      if (bl.npolproducts[f] <= 0 || bl.npolproducts[f] > FXG_MAX_POLPRODUCTS) return fail(FXG_EINVAL, "npolproducts");
      const fxg_datastream& d1 = j->datastreams[bl.ds1];
      const fxg_datastream& d2 = j->datastreams[bl.ds2];
      for (int p = 0; p < bl.npolproducts[f]; p++) {
        int b1 = bl.band1[f][p], b2 = bl.band2[f][p];
        if (b1 < 0 || b1 >= d1.nrecbands + d1.nzoombands || b2 < 0 || b2 >= d2.nrecbands + d2.nzoombands) return fail(FXG_EINVAL, "baseline band index");
        /* a band carries the baseline frequency or covers the same sky band
         * in the opposite sideband (the USB x LSB products of
         * configuration.cpp:2926-2937), with equal channel counts */
        auto carries = [&](int g) {
          const int fi = bl.freq_index[f];
          if (g == fi) return true;
          const fxg_freq& a = j->freqs[g]; const fxg_freq& c = j->freqs[fi];
          if ((a.lsb != 0) == (c.lsb != 0)) return false;
          const double la = a.freq_mhz - (a.lsb ? a.bw_mhz : 0.0), lc = c.freq_mhz - (c.lsb ? c.bw_mhz : 0.0);
          return la == lc && a.bw_mhz == c.bw_mhz && a.nchan == c.nchan && a.chans_to_avg == c.chans_to_avg;
        };
        if (!carries(ds_band_freqindex(d1, b1)) || !carries(ds_band_freqindex(d2, b2))) return fail(FXG_EINVAL, "baseline band frequency mismatch");
      }
    }
  }
  return 0;
}

/* populateResultLengths for one config, single phase centre, no pulsar
 * binning (configuration.cpp:2212-2441), plus the used / equivalent-used
 * frequency flags (configuration.cpp:325-367) and stride selection. */
/* phase-cal extractor kinds (PCal::getNew, pcal.cpp:139-203) */
enum { PC_DUMMY = 0, PC_TRIVIAL = 1, PC_SHIFT = 2, PC_IMPLICIT = 3,
       PC_CSHIFT = 4,        /* complex samples, mixer shift (PCalExtractorComplex) */
       PC_CIMPLICIT = 5 };   /* complex samples, implicit shift (PCalExtractorComplexImplicitShift) */
static void recfreq_pcal_tones(const fxg_freq& fq, double interval, double base, int* ntones, int* offsethz);
static int pcal_kind(double bw_hz, double spacing_hz, int offset_hz, double minres);
static int pcal_kind_c(double bw_hz, double spacing_hz, int offset_hz, double minres);
static int pcal_ntones_of(int kind, double bw_hz, double spacing_hz, int offset_hz);

static int make_layout(const fxg_job_desc* j, Layout* L) {
  int r = validate(j); if (r) return r;
  L->nfreqs = j->nfreqs; L->nds = j->ndatastreams; L->nbl = j->nbaselines;
  L->freqused.assign(j->nfreqs, 0); L->equivused.assign(j->nfreqs, 0);
  L->blfreq_local.assign(j->nbaselines, std::vector<int>(j->nfreqs, -1));
  int maxproducts = 0;
  for (int b = 0; b < j->nbaselines; b++)
    for (int f = 0; f < j->baselines[b].nfreqs; f++) {
      L->freqused[j->baselines[b].freq_index[f]] = 1;
      L->blfreq_local[b][j->baselines[b].freq_index[f]] = f;
      maxproducts = std::max(maxproducts, j->baselines[b].npolproducts[f]);
    }
  for (int a = 0; a < j->nfreqs; a++) {
    if (L->freqused[a]) continue;
    for (int k = 0; k < j->nfreqs; k++) {
      const fxg_freq& fa = j->freqs[a]; const fxg_freq& fk = j->freqs[k];
      double bwdiff = fa.bw_mhz - fk.bw_mhz;
      double freqdiff = fa.freq_mhz - fk.freq_mhz;
      if (fa.lsb) freqdiff -= fa.bw_mhz;
      if (fk.lsb) freqdiff += fk.bw_mhz;
      /* the reference compares signed differences against TINY (configuration.cpp:356-358) */
      if (bwdiff < TINY && freqdiff < TINY && fa.nchan == fk.nchan && fa.chans_to_avg == fk.chans_to_avg)
        if (L->freqused[k]) L->equivused[a] = 1;
    }
  }
  /* strides */
  L->arraystridelen.assign(j->ndatastreams, 0);
  for (int d = 0; d < j->ndatastreams; d++) {
    int nchan = j->freqs[j->datastreams[d].recfreq_index[0]].nchan;
    L->arraystridelen[d] = j->arraystridelen > 0 ? j->arraystridelen : calcstridelength(nchan);
  }
  if (j->xmacstridelen > 0) L->xmacstridelen = j->xmacstridelen;
  else {
    long nchangcd = 0;
    for (int b = 0; b < j->nbaselines; b++)
      for (int f = 0; f < j->baselines[b].nfreqs; f++) {
        long nc = j->freqs[j->baselines[b].freq_index[f]].nchan;
        nchangcd = nchangcd == 0 ? nc : gcdl(nchangcd, nc);
      }
    int target = 150, best = 0; double err = 1.0e12;
    for (long g = nchangcd; g > 0; --g) {
      if (nchangcd % g) continue;
      double dd = std::fabs((double)(g - target));
      if (dd < err) { err = dd; best = (int)g; }
    }
    L->xmacstridelen = best > 0 ? best : 1;
  }
  L->rotatestridelen = calcstridelength(L->xmacstridelen);
  for (int f = 0; f < j->nfreqs; f++)
    if (L->freqused[f] && j->freqs[f].nchan % L->xmacstridelen) return fail(FXG_EINVAL, "xmac stride must divide nchan");

  L->bandsperautocorr = maxproducts > 2 ? 2 : 1;
  int64_t idx = 0;
  L->bloff.assign(j->nfreqs, std::vector<int64_t>(j->nbaselines, -1));
  L->bwoff.assign(j->nfreqs, std::vector<int64_t>(j->nbaselines, -1));
  for (int f = 0; f < j->nfreqs; f++) {
    if (!L->freqused[f]) continue;
    for (int b = 0; b < j->nbaselines; b++) {
      int lf = L->blfreq_local[b][f];
      if (lf < 0) continue;
      L->bloff[f][b] = idx;
      idx += (int64_t)j->baselines[b].npolproducts[lf] * j->freqs[f].nchan / j->freqs[f].chans_to_avg;
    }
  }
  for (int f = 0; f < j->nfreqs; f++) {
    if (!L->freqused[f]) continue;
    for (int b = 0; b < j->nbaselines; b++) {
      int lf = L->blfreq_local[b][f];
      if (lf < 0) continue;
      L->bwoff[f][b] = idx;
      int np = j->baselines[b].npolproducts[lf];
      idx += np / 2 + np % 2;
    }
  }
  L->acoff.assign(j->ndatastreams, 0); L->acwoff.assign(j->ndatastreams, 0); L->pcaloff.assign(j->ndatastreams, 0);
  for (int d = 0; d < j->ndatastreams; d++) {
    const fxg_datastream& ds = j->datastreams[d];
    L->acoff[d] = idx;
    for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
      int fi = ds_band_freqindex(ds, k);
      if (L->freqused[fi] || L->equivused[fi])
        idx += (int64_t)L->bandsperautocorr * j->freqs[fi].nchan / j->freqs[fi].chans_to_avg;
    }
  }
  for (int d = 0; d < j->ndatastreams; d++) {
    const fxg_datastream& ds = j->datastreams[d];
    L->acwoff[d] = idx;
    int toadd = 0;
    for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
      int fi = ds_band_freqindex(ds, k);
      if (L->freqused[fi] || L->equivused[fi]) toadd += L->bandsperautocorr;
    }
    toadd = (toadd + 1) / 2;
    if (toadd == 0) toadd = 1;
    idx += toadd;
  }
  /* and finally the phase cal (configuration.cpp:2422-2431) */
  L->pcal_ntones.assign(j->ndatastreams, {}); L->pcal_offhz.assign(j->ndatastreams, {});
  for (int d = 0; d < j->ndatastreams; d++) {
    const fxg_datastream& ds = j->datastreams[d];
    L->pcaloff[d] = idx;
    if (!(ds.pcal_interval_mhz > 0)) continue;
    L->pcal_ntones[d].resize(ds.nrecfreqs); L->pcal_offhz[d].resize(ds.nrecfreqs);
    for (int i = 0; i < ds.nrecfreqs; i++)
      recfreq_pcal_tones(j->freqs[ds.recfreq_index[i]], ds.pcal_interval_mhz, ds.pcal_base_mhz,
                         &L->pcal_ntones[d][i], &L->pcal_offhz[d][i]);
    for (int k = 0; k < ds.nrecbands; k++) idx += L->pcal_ntones[d][ds.recband_freq[k]];
    /* Mode's check that configuration.cpp and pcal.cpp agree (mode.cpp:426-427) */
    const double bwhz = 1e6 * j->freqs[ds.recfreq_index[0]].bw_mhz;
    for (int i = 0; i < ds.nrecfreqs; i++) {
      if (L->pcal_offhz[d][i] < -1) return fail(FXG_EUNSUPPORTED, "phase cal: first tone outside the band");
      const int kind = ds.sampling == FXG_SAMPLING_COMPLEX
                           ? pcal_kind_c(bwhz, 1e6 * ds.pcal_interval_mhz, L->pcal_offhz[d][i], 1e6)
                           : pcal_kind(bwhz, 1e6 * ds.pcal_interval_mhz, L->pcal_offhz[d][i], 1e6);
      if (kind != PC_DUMMY && pcal_ntones_of(kind, bwhz, 1e6 * ds.pcal_interval_mhz, L->pcal_offhz[d][i]) !=
                                  L->pcal_ntones[d][i])
        return fail(FXG_EUNSUPPORTED, "phase cal: extractor and configuration tone counts differ");
    }
  }
  L->total = idx;
  return 0;
}

static int make_geom(const fxg_job_desc* j, const Layout& lay, int d, DsGeom* g) {
  const fxg_datastream& ds = j->datastreams[d];
  int r = make_fmt(&ds, &g->fmt); if (r) return r;
  const fxg_freq& f0 = j->freqs[ds.recfreq_index[0]];
  /* complex sampling: fftchannels = nchan, sample time 1/bw (mode.cpp:82-87) */
  g->cplx = g->fmt.cplx; g->dsb = g->cplx && ds.complex_dsb != 0;
  g->C = f0.nchan; g->L = g->cplx ? g->C : 2 * g->C; g->bw = f0.bw_mhz;
  g->s = lay.arraystridelen[d];
  if (g->L % g->s) return fail(FXG_EINVAL, "array stride must divide fftchannels");
  if (g->C % g->s) return fail(FXG_EINVAL, "array stride must divide nchan");
  g->numfrstrides = g->L / g->s; g->nfracsteps = g->C / g->s;
  g->sampletime = 1.0 / (2.0 * g->bw);
  if (g->cplx) g->sampletime *= 2.0;
  g->order = j->fringe_order;
  if (g->cplx && g->order == 0)
    return fail(FXG_EUNSUPPORTED, "post-F fringe rotation is not supported for complex data (mode.cpp:1155-1158)");
  g->fractionalLoFreq = false; g->deltapoloffsets = false;
  for (int i = 0; i < ds.nrecfreqs; i++) {
    const fxg_freq& fi = j->freqs[ds.recfreq_index[i]];
    double loval = fi.freq_mhz;
    if (g->dsb) loval += fi.lsb ? -fi.bw_mhz / 2.0 : fi.bw_mhz / 2.0;     /* mode.cpp:95-105 */
    if (std::fabs(loval - (int)(loval + 0.5)) > TINY) g->fractionalLoFreq = true;
    if (ds.recfreq_clockoffsetdelta_us[i] != 0.0) g->deltapoloffsets = true;
  }
  /* polarisation phase offsets (mode.cpp:337-346, 435-460, 1341-1353): when
   * any recorded frequency has one, each frequency with two or more bands
   * turns its LCP band (the second of its first two bands if the first is R,
   * else the first) by phasecorrA = cos(d) - i sin(d), d = -offset/2 degrees */
  g->polre.assign(ds.nrecbands, 1.0); g->polim.assign(ds.nrecbands, 0.0);
  bool phasepoloffset = false;
  for (int i = 0; i < ds.nrecfreqs; i++) phasepoloffset |= ds.recfreq_phaseoffset_deg[i] != 0.0;
  for (int i = 0; phasepoloffset && i < ds.nrecfreqs; i++) {
    int idx[2] = {-1, -1}, count = 0;
    for (int b = 0; b < ds.nrecbands; b++)
      if (ds.recband_freq[b] == i) { if (count < 2) idx[count] = b; count++; }
    if (count < 2) continue;
    const int lcp = ds.recband_pol[idx[0]] == 'R' ? idx[1] : idx[0];
    const double degphase = -ds.recfreq_phaseoffset_deg[i] / 2;
    g->polre[lcp] = std::cos(degphase * M_PI / 180.0);
    g->polim[lcp] = -std::sin(degphase * M_PI / 180.0);
  }
  if ((g->L & (g->L - 1)) != 0) return fail(FXG_EUNSUPPORTED, "fftchannels must be a power of two");
  /* Mk5Mode: blockclock = 2*bw, so samplesperblock = 1 (mode.cpp:129-141) */
  g->samplesperblock = 1;
  int numsamplebits = g->fmt.unitbits;
  g->bytesperblocknum = (ds.nrecbands * g->samplesperblock * numsamplebits) / 8;
  if (g->bytesperblocknum == 0) {
    g->bytesperblocknum = 1;
    g->bytesperblockden = 8 / (ds.nrecbands * g->samplesperblock * numsamplebits);
  } else g->bytesperblockden = 1;
  if (g->fmt.framebytes <= g->fmt.headerbytes) return fail(FXG_EINVAL, "framebytes");
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Mode::process restated (fringe orders 0, 1, 2, real sampling)           */
/* ------------------------------------------------------------------------ */
struct BlockGeom {
  int status; int nearestsample; int integerdelay; float weight; float fracsampleerror;
  int intwalltime; double a, b, c, fracwalltime, walltimesecs; double starttime, averagedelay;
};

static bool validbit(const fxg_station_input& in, int index) {
  if (!in.validflags || in.nvalidflags <= index / FXG_FLAGS_PER_INT) return false;
  return ((in.validflags[index / FXG_FLAGS_PER_INT] >> (index % FXG_FLAGS_PER_INT)) & 0x01) != 0;
}

static bool station_usable(const fxg_subint_desc& sub, const fxg_station_input& in) {
  if (!sub.model_ok) return false;               /* setOffsets !foundok -> INVALID_SUBINT */
  if (in.sec <= FXG_INVALID_SUBINT) return false;
  if (in.scan != sub.scan) return false;         /* mode.cpp:1532-1536 torches data */
  if (in.nbytes <= 1) return false;
  return true;
}

/* mode.cpp:686-772 (geometry, range check, unpack, zero-weight check) and
 * :789-838 (integer delay and interpolator for fringe orders 0, 1, 2).  The
 * unpacked samples go to `work`. */
static BlockGeom block_geom(const DsGeom& g, const fxg_subint_desc& sub,
                            const fxg_station_input& in, int index,
                            std::vector<float>* work, Unpacked* up) {
  BlockGeom bg; std::memset(&bg, 0, sizeof(bg));
  const int fftchannels = g.L;
  if (!station_usable(sub, in) || !validbit(in, index)) { bg.status = 1; return bg; }
  const double* interp = in.interpolator;
  double fftcentre = index + 0.5;
  double averagedelay = interp[0] * fftcentre * fftcentre + interp[1] * fftcentre + interp[2];
  double fftstartmicrosec = index * fftchannels * g.sampletime;
  double starttime = (sub.sec - in.sec) * 1000000.0 +
                     (static_cast<long long>(sub.ns) - static_cast<long long>(in.ns)) / 1000.0 +
                     fftstartmicrosec - averagedelay;
  int nearestsample = int(starttime / g.sampletime + 0.5);
  double walltimesecs = sub.scan_start_sec + sub.sec + sub.ns / 1.0e9 + fftstartmicrosec / 1.0e6;
  int intwalltime = static_cast<int>(walltimesecs);
  bg.averagedelay = averagedelay; bg.starttime = starttime; bg.walltimesecs = walltimesecs;
  bg.intwalltime = intwalltime; bg.fracwalltime = walltimesecs - intwalltime;
  bg.nearestsample = nearestsample;
  if (nearestsample < -1 ||
      ((((int64_t)nearestsample + fftchannels) / g.samplesperblock) * g.bytesperblocknum) / g.bytesperblockden > in.nbytes) {
    bg.status = 2; return bg;
  }
  if (nearestsample == -1) nearestsample = 0;
  bg.nearestsample = nearestsample;
  *up = mk5_unpack(g.fmt, in.data, in.nbytes, nearestsample, g.L, *work);
  bg.weight = up->weight;
  if (!(bg.weight > 0.0f)) { bg.status = 3; return bg; }
  double nearestsampletime = nearestsample * g.sampletime;
  bg.fracsampleerror = float(starttime - nearestsampletime);
  if (g.order == 0) {                                   /* post-F (mode.cpp:791-793) */
    bg.integerdelay = static_cast<int>(averagedelay);
    bg.a = 0.0; bg.b = 0.0; bg.c = averagedelay;
  } else if (g.order == 1) {                            /* linear (mode.cpp:794-813) */
    double d0 = interp[0] * index * index + interp[1] * index + interp[2];
    double d1 = interp[0] * (index + 0.5) * (index + 0.5) + interp[1] * (index + 0.5) + interp[2];
    double d2 = interp[0] * (index + 1) * (index + 1) + interp[1] * (index + 1) + interp[2];
    double a = d2 - d0;
    double b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
    int integerdelay = static_cast<int>(b);
    b -= integerdelay;
    bg.a = a; bg.b = b; bg.c = 0.0; bg.integerdelay = integerdelay;
  } else {                                              /* quadratic (mode.cpp:814-838) */
    double a = interp[0];
    double b = interp[1] + index * interp[0] * 2.0;
    double c = interp[2] + index * interp[1] + (index * index) * interp[0];
    int integerdelay = int(c);
    c -= integerdelay;
    bg.a = a; bg.b = b; bg.c = c; bg.integerdelay = integerdelay;
  }
  bg.status = 0;
  return bg;
}

template <typename T> struct StationWork {
  std::vector<float> unpacked;
  Unpacked up{};                    /* where the block's samples start in `unpacked` */
  std::vector<cpx<T>> rot, fracA, fracB, buf;
  std::vector<std::complex<T>> fftbuf;
};

/* One block of one station: spectra[nrecbands][C] (zeroed if bailed). */
template <typename T>
static BlockGeom station_block(const fxg_job_desc* j, const DsGeom& g, int d,
                               const fxg_subint_desc& sub, const fxg_station_input& in,
                               int index, const FFT<T>& fft, StationWork<T>& w,
                               cpx<T>* spectra /* [nrecbands*C] */, cpx<T>* prefrac = nullptr) {
  const fxg_datastream& ds = j->datastreams[d];
  const int C = g.C, L = g.L, s = g.s;
  Unpacked up{};
  BlockGeom bg = block_geom(g, sub, in, index, &w.unpacked, &up);
  w.up = up;
  std::memset((void*)spectra, 0, sizeof(cpx<T>) * (size_t)ds.nrecbands * C);
  if (prefrac) std::memset((void*)prefrac, 0, sizeof(cpx<T>) * (size_t)ds.nrecbands * C);
  if (bg.status != 0) return bg;
  const bool fp64 = sizeof(T) == 8;
  w.rot.resize(L); w.fracA.resize(C); w.fracB.resize(C); w.fftbuf.resize(L);
  /* x values of the sub and step rotator points (mode.cpp:804-812, 821-837) */
  std::vector<double> subxval(s), subquadxval(s), stepxval(g.numfrstrides);
  const double stepxoff1 = double(1 * s) / double(L);
  for (int k = 0; k < s; k++) {
    const double subxoff = double(k) / double(L);
    if (g.order == 1) subxval[k] = subxoff * bg.a + bg.b;
    else if (g.order == 2) {
      subxval[k] = subxoff * (bg.b + bg.a * stepxoff1) + bg.c;
      subquadxval[k] = subxoff * (2 * bg.a * stepxoff1);
    }
  }
  for (int jj = 0; jj < g.numfrstrides; jj++) {
    const double stepxoff = double(jj * s) / double(L);
    if (g.order == 1) stepxval[jj] = stepxoff * bg.a;
    else if (g.order == 2) stepxval[jj] = stepxoff * bg.b + (stepxoff * stepxoff) * bg.a;
  }

  for (int i = 0; i < ds.nrecfreqs; i++) {
    const fxg_freq& fq = j->freqs[ds.recfreq_index[i]];
    /* the LO of the fringe rotator (mode.cpp:886-908): the band edge; for
     * double-sideband complex data the band centre; negative for complex LSB */
    double lofreq = fq.freq_mhz;
    if (g.dsb) lofreq += fq.lsb ? -fq.bw_mhz / 2.0 : fq.bw_mhz / 2.0;
    if (g.cplx && fq.lsb) lofreq = -lofreq;
    const double looffset = ds.recfreq_looffset_hz[i];
    bool looff = false, isfraclooffset = false; double fraclooffset = 0.0;
    if (looffset > 0.0 || looffset < 0.0) {                                               /* mode.cpp:888-896 */
      looff = true;
      fraclooffset = std::fabs(looffset) - int(std::fabs(looffset));
      if (fraclooffset > TINY) isfraclooffset = true;
      if (looffset < 0) fraclooffset = -fraclooffset;
    }
    /* pre-F fringe rotator: order 1 (mode.cpp:909-978), order 2 (:980-1050);
     * order 0 rotates after the FFT (phase folded into the frac rotator) */
    if (g.order == 0) {
      for (int n = 0; n < L; n++) { w.rot[n].re = 1; w.rot[n].im = 0; }
    } else {
      std::vector<cpx<T>> sub_c(s), quad_c(s), step_c(g.numfrstrides);
      auto rot_of = [&](double ph, cpx<T>* out) {
        T arg = fp64 ? (T)(-TWO_PI * (ph - int(ph))) : (T)(float)(-TWO_PI * (ph - int(ph)));
        T sn, cs; sincos_t(arg, &sn, &cs); out->re = cs; out->im = sn;
      };
      for (int k = 0; k < s; k++) {
        double ph = subxval[k] * lofreq;
        if (g.fractionalLoFreq) ph += (lofreq - int(lofreq)) * double(bg.integerdelay);
        if (looff) {
          ph += (k * g.sampletime / 1e6) * (-looffset);
          ph += -looffset * bg.fracwalltime;
          if (isfraclooffset) ph += -fraclooffset * bg.intwalltime;
        }
        rot_of(ph, &sub_c[k]);
        if (g.order == 2) rot_of(subquadxval[k] * lofreq, &quad_c[k]);
      }
      for (int jj = 0; jj < g.numfrstrides; jj++) {
        double ph = stepxval[jj] * lofreq;
        if (looff) ph += (jj * s * g.sampletime / 1e6) * (-looffset);
        rot_of(ph, &step_c[jj]);
      }
      if (g.order == 1) {
        for (int k = 0; k < s; k++) w.rot[k] = sub_c[k];
        for (int jj = 1; jj < g.numfrstrides; jj++)
          for (int k = 0; k < s; k++) w.rot[(size_t)jj * s + k] = cmul(sub_c[k], step_c[jj]);
      } else {
        /* piecewise rotator: rot[j s + k] = piece_j[k] step[j], piece_0 = sub,
         * piece_{j+1} = piece_j * quad, in that operation order (mode.cpp:1043-1049) */
        std::vector<cpx<T>> piece(sub_c);
        for (int jj = 0; jj < g.numfrstrides; jj++)
          for (int k = 0; k < s; k++) {
            w.rot[(size_t)jj * s + k] = cmul(piece[k], step_c[jj]);
            piece[k] = cmul(piece[k], quad_c[k]);
          }
      }
    }

    /* fractional-sample rotator (mode.cpp:1054-1097, 1100-1146) */
    for (int pass = 0; pass < (g.deltapoloffsets ? 2 : 1); pass++) {
      cpx<T>* fr = pass == 0 ? w.fracA.data() : w.fracB.data();
      double dval = (double)(fp64 ? (double)(bg.starttime - bg.nearestsample * g.sampletime) : (double)bg.fracsampleerror)
                    - ds.recfreq_clockoffset_us[i]
                    + (pass == 0 ? 1.0 : -1.0) * ds.recfreq_clockoffsetdelta_us[i] / 2;
      T val = fp64 ? (T)dval : (T)(float)dval;
      std::vector<cpx<T>> stepf(g.nfracsteps);
      /* order 0: the fringe phase of the block, added to the sub arguments
       * (mode.cpp:1064-1073, 1111-1119) */
      T postf = 0;
      if (g.order == 0) {
        double phaserotation = (bg.averagedelay - bg.integerdelay) * lofreq;
        if (g.fractionalLoFreq) phaserotation += bg.integerdelay * (lofreq - int(lofreq));
        phaserotation -= bg.walltimesecs * looffset;
        postf = fp64 ? (T)(-TWO_PI * (phaserotation - int(phaserotation)))
                     : (T)(float)(-TWO_PI * (phaserotation - int(phaserotation)));
      }
      for (int k = 0; k < s; k++) {
        T cf = fp64 ? (T)((TWO_PI * (k) * g.bw) / C) : (T)(float)((TWO_PI * (k) * g.bw) / C);
        T arg = cf * val;
        if (g.order == 0) arg = arg + postf;
        T sn, cs; sincos_t(arg, &sn, &cs); fr[k].re = cs; fr[k].im = sn;
      }
      for (int jj = 0; jj < g.nfracsteps; jj++) {
        /* step channel frequencies (mode.cpp:328-334, 864-880): DSB, LSB, USB */
        double dcf = g.dsb ? (TWO_PI * jj * s * g.bw) / C - TWO_PI * g.bw / 2.0
                   : fq.lsb ? (-TWO_PI * ((g.nfracsteps - jj) * s) * g.bw) / C
                            : (TWO_PI * jj * s * g.bw) / C;
        T cf = fp64 ? (T)dcf : (T)(float)dcf;
        T arg = cf * val; T sn, cs; sincos_t(arg, &sn, &cs); stepf[jj].re = cs; stepf[jj].im = sn;
      }
      for (int jj = 1; jj < g.nfracsteps; jj++)
        for (int k = 0; k < s; k++) fr[(size_t)jj * s + k] = cmul(fr[k], stepf[jj]);
      for (int k = 0; k < s; k++) fr[k] = cmul(fr[k], stepf[0]);
    }
    if (!g.deltapoloffsets) std::copy(w.fracA.begin(), w.fracA.end(), w.fracB.begin());

    /* bands of this frequency (mode.cpp:1148-1310) */
    for (int jb = 0; jb < ds.nrecbands; jb++) {
      if (ds.recband_freq[jb] != i) continue;
      if (g.cplx) {
        /* rotator times the complex samples (vectorMul_cf32, mode.cpp:1184-1189) */
        const float* x = w.unpacked.data() + 2 * ((size_t)jb * up.stride + up.munged);
        for (int n = 0; n < L; n++) {
          T xr = (T)x[2 * n], xi = (T)x[2 * n + 1];
          w.fftbuf[n] = std::complex<T>(w.rot[n].re * xr - w.rot[n].im * xi, w.rot[n].re * xi + w.rot[n].im * xr);
        }
      } else {
        const float* x = w.unpacked.data() + (size_t)jb * up.stride + up.munged;
        for (int n = 0; n < L; n++) {
          T xv = (T)x[n];
          w.fftbuf[n] = std::complex<T>(xv * w.rot[n].re, xv * w.rot[n].im);
        }
      }
      fft.run(w.fftbuf.data());
      const int off = fq.lsb ? C : 0;
      const cpx<T>* fr = (!g.deltapoloffsets || ds.recband_pol[jb] == 'R') ? w.fracA.data() : w.fracB.data();
      cpx<T>* out = spectra + (size_t)jb * C;
      for (int c = 0; c < C; c++) {
        cpx<T> v;
        if (g.cplx) {
          /* channel c of the band (mode.cpp:1208-1241): USB as transformed,
           * DSB with the halves swapped; LSB conjugated in reversed order
           * (vectorConjFlip_cf32), DC kept in channel 0 without conjugation
           * for single sideband */
          int m = c; bool cj = false;
          if (!fq.lsb) m = g.dsb ? (c + C / 2) % C : c;
          else if (g.dsb) { m = ((C / 2 - c) % C + C) % C; cj = true; }
          else if (c > 0) { m = C - c; cj = true; }
          v.re = w.fftbuf[m].real(); v.im = cj ? -w.fftbuf[m].imag() : w.fftbuf[m].imag();
        } else {
          v.re = w.fftbuf[off + c].real(); v.im = w.fftbuf[off + c].imag();
        }
        if (prefrac) prefrac[(size_t)jb * C + c] = v;     /* fftoutputs before the frac rotation */
        cpx<T> o; o.re = v.re * fr[c].re - v.im * fr[c].im; o.im = v.re * fr[c].im + v.im * fr[c].re;
        if (g.polre[jb] != 1.0 || g.polim[jb] != 0.0) {   /* vectorMulC_cf32_I(phasecorrA) */
          const T pr = (T)g.polre[jb], pi = (T)g.polim[jb];
          const T re = o.re * pr - o.im * pi;
          o.im = o.re * pi + o.im * pr;
          o.re = re;
        }
        out[c] = o;
      }
    }
  }
  return bg;
}

/* acc += a * conj(b), the operation order of genericAddProduct_32fc applied
 * to (a, conj(b)) (architecture.h.in:580-586). */
template <typename T, typename A>
static inline void addprod_conj(A* acc, const cpx<T>& a, const cpx<T>& b) {
  T cre = b.re, cim = -b.im;
  acc[0] += a.re * cre - a.im * cim;
  acc[1] += a.re * cim + a.im * cre;
}

/* ------------------------------------------------------------------------ */
/* Phase cal                                                                */
/* ------------------------------------------------------------------------ */
/* mathutil.cpp:8-30 */
static long long gcd_l(long a, long b) {
  if (a == 0 || b == 0) return std::max(a, b);
  while (true) {
    a = a % b;
    if (a == 0) return b;
    b = b % a;
    if (b == 0) return a;
  }
}
static long long gcd_d(double a, double b) { return gcd_l((long)std::floor(a + 0.5), (long)std::floor(b + 0.5)); }

/* PCal::calcNumTones (pcal.cpp:212-229) */
static int calc_num_tones(double bw, double offset, double step) {
  int n = 0;
  if (bw <= 0.0f || step <= 0.0f) return 0;
  double ftone = offset;
  while (ftone <= bw) {
    n++;
    ftone += step;
    if (step <= 0.0f) break;
  }
  if (offset == 0.0 && ftone - step == bw) n--;
  return n;
}

/* Tones of one recorded frequency (configuration.cpp:1771-1845) */
static void recfreq_pcal_tones(const fxg_freq& fq, double interval, double base, int* ntones, int* offsethz) {
  int n = 0;
  const double lofreq = fq.freq_mhz, bw = fq.bw_mhz;
  double tonefreq;
  if (fq.lsb) {
    tonefreq = double(int(lofreq / interval)) * interval;
    if (tonefreq == lofreq) tonefreq -= interval;
    if (tonefreq >= lofreq - bw)
      while (tonefreq - n * interval >= lofreq - bw) n++;
    if (n > 0) {
      double t0 = 1e6 * tonefreq - double(0) * 1e6 * interval - 1e6 * base;
      *offsethz = (int)long(1e6 * lofreq - t0 + 0.5);
    }
  } else {
    tonefreq = double(int(lofreq / interval)) * interval;
    if (tonefreq <= lofreq) tonefreq += interval;
    if (tonefreq <= lofreq + bw)
      while (tonefreq + n * interval <= lofreq + bw) n++;
    if (n > 0) {
      double t0 = 1e6 * tonefreq + 0 * 1e6 * interval + 1e6 * base;
      *offsethz = (int)long(t0 - 1e6 * lofreq + 0.5);
    }
  }
  if (n == 0) { n = 1; *offsethz = -1; }
  *ntones = n;
}

/* Forward DFT bin k of x[N] (vectorDFT_CtoC_cf32, unnormalised), in FP64 */
template <typename T>
static cpx<double> dft_bin(const cpx<T>* x, long N, long k) {
  double re = 0, im = 0;
  for (long n = 0; n < N; n++) {
    const long m = (long)(((long long)n * k) % N);
    const double a = -TWO_PI * double(m) / double(N);
    const double c = std::cos(a), sn = std::sin(a);
    re += double(x[n].re) * c - double(x[n].im) * sn;
    im += double(x[n].re) * sn + double(x[n].im) * c;
  }
  cpx<double> o; o.re = re; o.im = im; return o;
}

/* The extractors of pcal.cpp, with their state: for real samples trivial
 * (:333-470), shifting (:498-690), implicit shift (:927-1104); for complex
 * samples shifting (:707-925) and implicit shift (:1107-1285); dummy
 * (:1376-1450).  T is the accumulator type (f32 in the reference).  The
 * complex kinds take len complex samples (float pairs, real first), sample at
 * the band's rate fs = bandwidth. */
template <typename T>
struct PcalExtractor {
  int kind = PC_DUMMY;
  double fs = 0, spacing = 0;
  long offset = 0;
  int nbins = 0, ntones = 0, tonestep = 0;
  bool dsb = false;                     /* complex double sideband: the DFT read reversed */
  size_t rotatorlen = 0, rotator_index = 0, pcal_index = 0;
  uint64_t samplecount = 0;
  bool finalized = false;
  std::vector<T> preal;
  std::vector<cpx<T>> pcplx, rotated, rotator;

  bool complex_kind() const { return kind == PC_CSHIFT || kind == PC_CIMPLICIT; }
  void make_rotator() {                 /* f32(cos(arg)), f32(sin(arg)) over 2 x rotatorlen */
    rotator.resize(2 * rotatorlen);
    const double dphi = 2 * M_PI * (-offset / fs);
    for (size_t n = 0; n < 2 * rotatorlen; n++) {
      const double arg = dphi * double(n);
      rotator[n].re = (T)std::cos(arg);
      rotator[n].im = (T)std::sin(arg);
    }
    rotated.assign(2 * rotatorlen, cpx<T>{0, 0});
  }
  void init(int k, double bw_hz, double spacing_hz, int offset_hz, double minres, bool dsb_ = false) {
    kind = k; spacing = std::fabs(spacing_hz); offset = offset_hz; fs = 2 * bw_hz; dsb = false;
    if (complex_kind()) {
      fs = bw_hz;
      dsb = dsb_;
      ntones = calc_num_tones(bw_hz, (double)offset, spacing);
      if (kind == PC_CIMPLICIT) {         /* pcal.cpp:1107-1128 */
        nbins = (int)(fs / gcd_d(fs, (double)offset));
        while (fs / (2 * nbins) > minres) nbins *= 2;
      } else {                            /* pcal.cpp:707-760 */
        nbins = (int)(fs / gcd_d(fs, spacing));
        while (fs / (2 * nbins) > minres) nbins *= 2;
        rotatorlen = (size_t)(fs / gcd_d(fs, (double)offset));
        rotatorlen = nbins * rotatorlen / gcd_l((long)rotatorlen, (long)nbins);
        make_rotator();
      }
    } else if (kind == PC_TRIVIAL) {
      offset = 0;
      nbins = (int)(fs / gcd_d(fs, spacing));
      ntones = calc_num_tones(bw_hz, 0.0f, spacing) - 1;
      tonestep = (int)(spacing / gcd_d(fs, spacing));
    } else if (kind == PC_SHIFT) {
      nbins = (int)(fs / gcd_d(fs, spacing));
      ntones = calc_num_tones(bw_hz, offset, spacing);
      rotatorlen = (size_t)(fs / gcd_d(fs, (double)offset));
      rotatorlen = nbins * rotatorlen / gcd_l((long)rotatorlen, (long)nbins);
      make_rotator();
    } else if (kind == PC_IMPLICIT) {
      nbins = (int)(fs / gcd_d(fs, (double)offset));
      ntones = calc_num_tones(bw_hz, offset, spacing);
      while (fs / (2 * nbins) > minres) nbins *= 2;
    } else {
      ntones = calc_num_tones(bw_hz, (double)offset, spacing);
      nbins = 2 * ntones;
    }
    clear();
  }
  void clear() {
    samplecount = 0; finalized = false;
    pcplx.assign(2 * (size_t)std::max(nbins, 1), cpx<T>{0, 0});
    preal.assign(2 * (size_t)std::max(nbins, 1), (T)0);
    if (kind == PC_SHIFT || kind == PC_CSHIFT) rotated.assign(2 * rotatorlen, cpx<T>{0, 0});
  }
  void adjust(size_t so) {
    if (kind == PC_SHIFT || kind == PC_CSHIFT) { rotator_index = so % rotatorlen; pcal_index = so % nbins; }
    else if (kind == PC_TRIVIAL || kind == PC_IMPLICIT) { rotator_index = 0; pcal_index = so % nbins; }
    else if (kind == PC_CIMPLICIT) pcal_index = so % nbins;
  }
  /* complex samples: src holds len (re, im) pairs */
  void extract_c(const float* src_f, size_t len) {
    if (finalized) return;
    const cpx<float>* src = reinterpret_cast<const cpx<float>*>(src_f);
    cpx<T>* dst = &pcplx[pcal_index];
    auto add = [&](const cpx<T>* p, size_t n) {      /* vectorAdd_cf32_I */
      for (size_t i = 0; i < n; i++) { dst[i].re += p[i].re; dst[i].im += p[i].im; }
    };
    if (kind == PC_CIMPLICIT) {           /* pcal.cpp:1195-1226 */
      std::vector<cpx<T>> x(nbins);
      auto take = [&](size_t n) { for (size_t i = 0; i < n; i++) x[i] = cpx<T>{(T)src[i].re, (T)src[i].im}; };
      const size_t tail = len % nbins, end = len - tail;
      for (size_t n = 0; n < end; n += nbins, src += nbins) { take(nbins); add(x.data(), nbins); }
      if (tail != 0) {
        take(tail); add(x.data(), tail);
        pcal_index = (pcal_index + tail) % nbins;
      }
    } else if (kind == PC_CSHIFT) {       /* pcal.cpp:836-878 */
      const size_t tail = len % rotatorlen, end = len - tail;
      auto mix = [&](const cpx<float>* x, size_t n) {  /* vectorMul_cf32 */
        for (size_t i = 0; i < n; i++) {
          const cpx<T> r = rotator[rotator_index + i];
          rotated[rotator_index + i].re = (T)x[i].re * r.re - (T)x[i].im * r.im;
          rotated[rotator_index + i].im = (T)x[i].re * r.im + (T)x[i].im * r.re;
        }
      };
      for (size_t n = 0; n < end; n += rotatorlen, src += rotatorlen) {
        mix(src, rotatorlen);
        const cpx<T>* pulse = &rotated[rotator_index];
        for (size_t p = 0; p < rotatorlen / nbins; p++) { add(pulse, nbins); pulse += nbins; }
      }
      if (tail != 0) {
        mix(src, tail);
        const cpx<T>* pulse = &rotated[rotator_index];
        const size_t tail2 = tail % nbins, end2 = tail - tail2;
        for (size_t p = 0; p < end2 / (size_t)nbins; p++) { add(pulse, nbins); pulse += nbins; }
        add(pulse, tail2);
        rotator_index = (rotator_index + tail) % rotatorlen;
        pcal_index = (pcal_index + tail2) % nbins;
      }
    }
    samplecount += len;
  }
  void extract(const float* src, size_t len) {
    if (finalized) return;
    if (complex_kind()) { extract_c(src, len); return; }
    if (kind == PC_TRIVIAL || kind == PC_IMPLICIT) {
      T* dst = &preal[pcal_index];
      const size_t tail = len % nbins, end = len - tail;
      for (size_t n = 0; n < end; n += nbins, src += nbins)
        for (int i = 0; i < nbins; i++) dst[i] += (T)src[i];
      if (tail != 0) {
        for (size_t i = 0; i < tail; i++) dst[i] += (T)src[i];
        pcal_index = (pcal_index + tail) % nbins;
      }
    } else if (kind == PC_SHIFT) {
      cpx<T>* dst = &pcplx[pcal_index];
      const size_t tail = len % rotatorlen, end = len - tail;
      auto mix = [&](const float* x, size_t n) {       /* vectorMul_f32cf32 */
        for (size_t i = 0; i < n; i++) {
          rotated[rotator_index + i].re = (T)x[i] * rotator[rotator_index + i].re;
          rotated[rotator_index + i].im = (T)x[i] * rotator[rotator_index + i].im;
        }
      };
      auto add = [&](const cpx<T>* p, size_t n) {
        for (size_t i = 0; i < n; i++) { dst[i].re += p[i].re; dst[i].im += p[i].im; }
      };
      for (size_t n = 0; n < end; n += rotatorlen, src += rotatorlen) {
        mix(src, rotatorlen);
        const cpx<T>* pulse = &rotated[rotator_index];
        for (size_t p = 0; p < rotatorlen / nbins; p++) { add(pulse, nbins); pulse += nbins; }
      }
      if (tail != 0) {
        mix(src, tail);
        const cpx<T>* pulse = &rotated[rotator_index];
        const size_t tail2 = tail % nbins, end2 = tail - tail2;
        for (size_t p = 0; p < end2 / (size_t)nbins; p++) { add(pulse, nbins); pulse += nbins; }
        add(pulse, tail2);
        rotator_index = (rotator_index + tail) % rotatorlen;
        pcal_index = (pcal_index + tail2) % nbins;
      }
    }
    samplecount += len;
  }
  /* getFinalPCal: out[ntones] */
  uint64_t final(cpx<T>* out) {
    if (kind == PC_DUMMY) {
      finalized = true;
      for (int i = 0; i < ntones; i++) out[i] = cpx<T>{0, 0};
      return samplecount;
    }
    if (!finalized) {
      finalized = true;
      if (kind == PC_TRIVIAL) {
        for (int i = 0; i < nbins; i++) preal[i] += preal[nbins + i];
        for (int i = 0; i < nbins; i++) { pcplx[i].re = preal[i]; pcplx[i].im = 0; }
      } else if (kind == PC_IMPLICIT) {
        for (int i = 0; i < 2 * nbins; i++) { pcplx[i].re = preal[i]; pcplx[i].im = 0; }
        for (int i = 0; i < nbins; i++) { pcplx[i].re += pcplx[nbins + i].re; pcplx[i].im += pcplx[nbins + i].im; }
      } else {
        for (int i = 0; i < nbins; i++) { pcplx[i].re += pcplx[nbins + i].re; pcplx[i].im += pcplx[nbins + i].im; }
      }
    }
    auto bin = [&](long k) { cpx<double> d = dft_bin(pcplx.data(), nbins, k); return cpx<T>{(T)d.re, (T)d.im}; };
    if (complex_kind()) {
      /* pcal.cpp:906-923 and 1262-1283: double sideband reverses the DFT
       * output (n <-> N-1-n); tone n at bin [offset bin +] n x step, the
       * copying stops at the first bin past the DFT (zeros here; the
       * reference leaves those tones as they were) */
      const size_t step = (size_t)(std::floor(double(nbins) * (spacing / fs)));
      const size_t off = kind == PC_CIMPLICIT ? (size_t)(std::floor(double(nbins) * (offset / fs))) : 0;
      for (size_t n = 0; n < (size_t)ntones; n++) {
        const size_t idx = off + n * step;
        if (idx >= (size_t)nbins) { out[n] = cpx<T>{0, 0}; continue; }
        out[n] = bin(dsb ? (long)(nbins - 1 - idx) : (long)idx);
      }
    } else if (kind == PC_TRIVIAL) {
      for (int i = 0; i < ntones; i++) out[i] = bin((long)(i + 1) * tonestep);
    } else if (kind == PC_IMPLICIT) {
      const size_t step = (size_t)(std::floor(double(nbins) * (spacing / fs)));
      const size_t off = (size_t)(std::floor(double(nbins) * (offset / fs)));
      for (size_t n = 0; n < (size_t)ntones; n++) {
        const size_t idx = off + n * step;
        out[n] = idx >= (size_t)nbins ? cpx<T>{0, 0} : bin((long)idx);
      }
    } else {
      const size_t step = (size_t)(std::floor(nbins * (spacing / fs)));
      for (size_t n = 0; n < (size_t)ntones; n++) {
        const size_t idx = n * step;
        if (idx >= (size_t)nbins) { out[n] = cpx<T>{0, 0}; continue; }   /* left untouched by the reference */
        out[n] = bin((long)idx);
      }
    }
    return samplecount;
  }
};

/* PCal::getNew for real data (pcal.cpp:139-203) */
static int pcal_kind(double bw_hz, double spacing_hz, int offset_hz, double minres) {
  const double fs = 2.0 * bw_hz;
  int No, Np, Nt;
  if (offset_hz == -1) { No = Np = Nt = 0; }
  else {
    No = (int)(fs / gcd_d(fs, (double)offset_hz));
    Np = (int)(fs / gcd_d(fs, spacing_hz));
    Nt = calc_num_tones(bw_hz, (double)offset_hz, spacing_hz);
    while (fs / (2 * No) > minres) No *= 2;
  }
  if (Nt == 0) return PC_DUMMY;
  if (offset_hz == 0.0f) return PC_TRIVIAL;
  if ((No % Np) == 0) return PC_IMPLICIT;
  return PC_SHIFT;
}

/* PCal::getNew for complex data (pcal.cpp:139-168, 186-191): the sample
 * rate is the bandwidth; the same period test picks the implicit shift */
static int pcal_kind_c(double bw_hz, double spacing_hz, int offset_hz, double minres) {
  const double fs = bw_hz;
  if (offset_hz == -1) return PC_DUMMY;
  long No = (long)(fs / gcd_d(fs, (double)offset_hz));
  const long Np = (long)(fs / gcd_d(fs, spacing_hz));
  const int Nt = calc_num_tones(bw_hz, (double)offset_hz, spacing_hz);
  while (fs / (2 * No) > minres) No *= 2;
  if (Nt == 0) return PC_DUMMY;
  return (No % Np) == 0 ? PC_CIMPLICIT : PC_CSHIFT;
}

static int pcal_ntones_of(int kind, double bw_hz, double spacing_hz, int offset_hz) {
  if (kind == PC_TRIVIAL) return calc_num_tones(bw_hz, 0.0f, spacing_hz) - 1;
  return calc_num_tones(bw_hz, (double)offset_hz, spacing_hz);
}

/* PCal::extractAndIntegrate_reference (pcal.cpp:253-331): the slow direct
 * method the reference keeps for comparison.  out[ntones]; state is the
 * running sample count. */
static void pcal_direct(double fs, long offset, int ntones, uint64_t* samplecount, const float* data, size_t len,
                        uint64_t sampleoffset, cpx<double>* out) {
  const size_t Nbins = 2 * ntones;
  const size_t maxtoneperiod = (size_t)(fs / gcd_l(offset, (long)fs));
  const double dphi = 2 * M_PI * (-offset / fs);
  std::vector<cpx<float>> acc(Nbins, cpx<float>{0, 0});
  for (size_t n = 0; n < len; n++) {
    const size_t no = n + sampleoffset;
    const int b = (int)(no % Nbins);
    const double phi = dphi * ((*samplecount + no) % maxtoneperiod);
    acc[b].re += std::cos(phi) * data[n];
    acc[b].im += std::sin(phi) * data[n];
  }
  *samplecount += len;
  for (int i = 0; i < ntones; i++) out[i] = dft_bin(acc.data(), (long)Nbins, offset == 0 ? i + 1 : i);
}

/* ------------------------------------------------------------------------ */
/* Core::processdata restated                                               */
/* ------------------------------------------------------------------------ */
template <typename T>
struct CoreRun {
  const fxg_job_desc* j; Layout lay; std::vector<DsGeom> geom;
  int init() {
    int r = make_layout(j, &lay); if (r) return r;
    geom.resize(j->ndatastreams);
    for (int d = 0; d < j->ndatastreams; d++) { r = make_geom(j, lay, d, &geom[d]); if (r) return r; }
    return 0;
  }
  /* Accumulate blocks [b0, b1) into acc (result layout, T precision). */
  void run_range(const fxg_subint_desc& sub, const fxg_station_input* st, int b0, int b1, std::vector<T>& acc) {
    const int nds = j->ndatastreams;
    std::vector<std::vector<cpx<T>>> spec(nds);
    std::vector<StationWork<T>> work(nds);
    std::vector<std::unique_ptr<FFT<T>>> ffts(nds);
    std::vector<std::vector<T>> ac0(nds), ac1(nds);      /* autocorrelations[2][band][C] as re,im */
    std::vector<std::vector<T>> w0(nds), w1(nds);
    for (int d = 0; d < nds; d++) {
      const fxg_datastream& ds = j->datastreams[d];
      spec[d].resize((size_t)ds.nrecbands * geom[d].C);
      ffts[d].reset(new FFT<T>(geom[d].L));
      ac0[d].assign((size_t)ds.nrecbands * geom[d].C * 2, 0); ac1[d].assign((size_t)ds.nrecbands * geom[d].C * 2, 0);
      w0[d].assign(ds.nrecbands, 0); w1[d].assign(ds.nrecbands, 0);
    }
    std::vector<float> dweight(nds);
    /* getDataWeight per (datastream, recorded band): the INTERLACEDVDIF
     * per-band weights, else the block's weight (mode.h:168) */
    std::vector<std::vector<float>> bweight(nds);
    std::vector<char> edv4(nds, 0);
    for (int d = 0; d < nds; d++) {
      bweight[d].assign((size_t)j->datastreams[d].nrecbands, 0.0f);
      edv4[d] = geom[d].fmt.edv4 ? 1 : 0;
    }
    /* this thread's phase-cal extractors per (datastream, recorded band),
     * created as Mode's constructor does (mode.cpp:412-428) */
    std::vector<std::vector<PcalExtractor<T>>> pcx(nds);
    std::vector<int> datasamples(nds, 0);
    for (int d = 0; d < nds; d++) {
      const fxg_datastream& ds = j->datastreams[d];
      if (lay.pcal_ntones[d].empty()) continue;
      const double bwhz = 1e6 * j->freqs[ds.recfreq_index[0]].bw_mhz;
      pcx[d].resize(ds.nrecbands);
      for (int k = 0; k < ds.nrecbands; k++) {
        const int off = lay.pcal_offhz[d][ds.recband_freq[k]];
        const int kind = geom[d].cplx ? pcal_kind_c(bwhz, 1e6 * ds.pcal_interval_mhz, off, 1e6)
                                      : pcal_kind(bwhz, 1e6 * ds.pcal_interval_mhz, off, 1e6);
        pcx[d][k].init(kind, bwhz, 1e6 * ds.pcal_interval_mhz, off, 1e6, geom[d].dsb);
      }
      datasamples[d] = static_cast<int>(st[d].ns / (geom[d].sampletime * 1e3) + 0.5);   /* mode.cpp:1572 */
    }
    /* this thread's cross products before averaging (threadcrosscorrs):
     * [freq][baseline][product][nchan] at xoff[f][b] */
    std::vector<std::vector<int64_t>> xoff(j->nfreqs, std::vector<int64_t>(j->nbaselines, -1));
    int64_t xlen = 0;
    for (int f = 0; f < j->nfreqs; f++) {
      if (!lay.freqused[f]) continue;
      for (int b = 0; b < j->nbaselines; b++) {
        int lf = lay.blfreq_local[b][f]; if (lf < 0) continue;
        xoff[f][b] = xlen; xlen += (int64_t)j->baselines[b].npolproducts[lf] * j->freqs[f].nchan;
      }
    }
    std::vector<T> tx(2 * (size_t)xlen, (T)0);
    for (int blk = b0; blk < b1; blk++) {
      for (int d = 0; d < nds; d++) {
        const fxg_datastream& ds = j->datastreams[d];
        BlockGeom bg = station_block<T>(j, geom[d], d, sub, st[d], blk, *ffts[d], work[d], spec[d].data());
        dweight[d] = bg.status == 0 ? bg.weight : 0.0f;
        for (int jb = 0; jb < ds.nrecbands; jb++)
          bweight[d][(size_t)jb] = bg.status != 0 ? 0.0f : edv4[d] ? work[d].up.bandweight[(size_t)jb] : bg.weight;
        if (bg.status != 0) continue;
        const int C = geom[d].C;
        /* phase cal of the block's L samples (mode.cpp:774-788) */
        for (size_t k = 0; k < pcx[d].size(); k++) {
          pcx[d][k].adjust((size_t)(datasamples[d] + bg.nearestsample));
          const size_t at = (size_t)k * work[d].up.stride + work[d].up.munged;   /* complex: float pairs */
          pcx[d][k].extract(work[d].unpacked.data() + (geom[d].cplx ? 2 * at : at), (size_t)geom[d].L);
        }
        for (int i = 0; i < ds.nrecfreqs; i++) {
          int idx[FXG_MAX_BANDS]; int count = 0;
          for (int jb = 0; jb < ds.nrecbands; jb++) {
            if (ds.recband_freq[jb] != i) continue;
            idx[count++] = jb;
            const cpx<T>* X = spec[d].data() + (size_t)jb * C;
            T* a0 = ac0[d].data() + (size_t)jb * C * 2;
            for (int c = 0; c < C; c++) addprod_conj<T, T>(a0 + 2 * c, X[c], X[c]);
            w0[d][jb] += (T)bweight[d][(size_t)jb];                  /* mode.cpp:1300-1307 */
          }
          if (count > 1 && j->crosspol_autocorrs) {            /* mode.cpp:1356-1375 */
            const cpx<T>* X0 = spec[d].data() + (size_t)idx[0] * C;
            const cpx<T>* X1 = spec[d].data() + (size_t)idx[1] * C;
            T* a10 = ac1[d].data() + (size_t)idx[0] * C * 2;
            T* a11 = ac1[d].data() + (size_t)idx[1] * C * 2;
            for (int c = 0; c < C; c++) { addprod_conj<T, T>(a10 + 2 * c, X0[c], X1[c]); addprod_conj<T, T>(a11 + 2 * c, X1[c], X0[c]); }
            const float wx = edv4[d] ? bweight[d][(size_t)idx[0]] * bweight[d][(size_t)idx[1]] : bg.weight;
            w1[d][idx[0]] += (T)wx; w1[d][idx[1]] += (T)wx;          /* mode.cpp:1365-1373 */
          }
        }
      }
      /* XMAC + baseline weights (core.cpp:826-975, 998-1045) */
      for (int f = 0; f < j->nfreqs; f++) {
        if (!lay.freqused[f]) continue;
        const int nchan = j->freqs[f].nchan;
        for (int b = 0; b < j->nbaselines; b++) {
          int lf = lay.blfreq_local[b][f]; if (lf < 0) continue;
          const fxg_baseline& bl = j->baselines[b];
          for (int p = 0; p < bl.npolproducts[lf]; p++) {
            const cpx<T>* v1 = band_spectrum(bl.ds1, bl.band1[lf][p], spec);
            const cpx<T>* v2 = band_spectrum(bl.ds2, bl.band2[lf][p], spec);
            T* out = tx.data() + 2 * (size_t)(xoff[f][b] + (int64_t)p * nchan);
            for (int c = 0; c < nchan; c++) addprod_conj<T, T>(out + 2 * c, v1[c], v2[c]);
            T* bw = acc.data() + 2 * (size_t)lay.bwoff[f][b] + p;
            /* getDataWeight of each product's recorded (parent) band (core.cpp:1028-1037) */
            const float wa = edv4[bl.ds1] ? bweight[bl.ds1][(size_t)record_band(bl.ds1, bl.band1[lf][p])] : dweight[bl.ds1];
            const float wb = edv4[bl.ds2] ? bweight[bl.ds2][(size_t)record_band(bl.ds2, bl.band2[lf][p])] : dweight[bl.ds2];
            *bw += (T)(wa * wb);
          }
        }
      }
    }
    /* uvshiftAndAverageBaselineFreq, single phase centre (core.cpp:1586-1802):
     * per xmac stride, copy, or average groups of channelinc channels with
     * vectorMean and add mean / stridestoaverage */
    const int xmacstridelen = lay.xmacstridelen;
    for (int f = 0; f < j->nfreqs; f++) {
      if (!lay.freqused[f]) continue;
      const int nchan = j->freqs[f].nchan, channelinc = j->freqs[f].chans_to_avg;
      int stridestoaverage = channelinc / xmacstridelen; if (stridestoaverage == 0) stridestoaverage = 1;
      int averagesperstride = xmacstridelen / channelinc; if (averagesperstride == 0) averagesperstride = 1;
      const int averagelength = xmacstridelen / averagesperstride;
      for (int b = 0; b < j->nbaselines; b++) {
        int lf = lay.blfreq_local[b][f]; if (lf < 0) continue;
        for (int p = 0; p < j->baselines[b].npolproducts[lf]; p++) {
          for (int x = 0; x < nchan / xmacstridelen; x++) {
            const T* src = tx.data() + 2 * (size_t)(xoff[f][b] + (int64_t)p * nchan + (int64_t)x * xmacstridelen);
            int64_t dest = lay.bloff[f][b] + ((int64_t)p * nchan + (int64_t)x * xmacstridelen) / channelinc;
            if (channelinc == 1) {
              for (int c = 0; c < 2 * xmacstridelen; c++) acc[2 * (size_t)dest + c] += src[c];
            } else {
              for (int l = 0; l < averagesperstride; l++) {
                T mre = 0, mim = 0;                          /* genericMean_32fc */
                for (int i = 0; i < averagelength; i++) { mre += src[2 * (l * averagelength + i)]; mim += src[2 * (l * averagelength + i) + 1]; }
                mre /= averagelength; mim /= averagelength;
                acc[2 * (size_t)dest] += mre / stridestoaverage;
                acc[2 * (size_t)dest + 1] += mim / stridestoaverage;
                dest++;
              }
            }
          }
        }
      }
    }
    /* Mode::averageFrequency with the datastream's channelstoaverage (its
     * first recorded frequency's, configuration.cpp:962), then
     * averageAndSendAutocorrs (core.cpp:1253-1369) */
    for (int d = 0; d < nds; d++) {
      const fxg_datastream& ds = j->datastreams[d];
      const int savg = j->freqs[ds.recfreq_index[0]].chans_to_avg;
      if (savg > 1) {
        for (std::vector<T>* ac : {&ac0[d], &ac1[d]})
          for (int jb = 0; jb < ds.nrecbands; jb++) {
            T* a = ac->data() + (size_t)jb * geom[d].C * 2;
            for (int k = 0; k < geom[d].C / savg; k++) {
              T mre = 0, mim = 0;
              for (int i = 0; i < savg; i++) { mre += a[2 * (k * savg + i)]; mim += a[2 * (k * savg + i) + 1]; }
              a[2 * k] = mre / savg; a[2 * k + 1] = mim / savg;
            }
          }
      }
      int64_t ri = lay.acoff[d];
      for (int pass = 0; pass < lay.bandsperautocorr; pass++) {
        if (pass == 1 && !j->crosspol_autocorrs) break;
        for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
          int fi = ds_band_freqindex(ds, k);
          if (!(lay.freqused[fi] || lay.equivused[fi])) continue;
          int nch = j->freqs[fi].nchan / j->freqs[fi].chans_to_avg;
          int parent = k < ds.nrecbands ? k : ds.zoomband_parent[k - ds.nrecbands];
          int choff = k < ds.nrecbands ? 0 : ds.zoomband_choffset[k - ds.nrecbands] / savg;   /* mode.cpp:396 */
          const T* src = (pass == 0 ? ac0[d] : ac1[d]).data() + ((size_t)parent * geom[d].C + choff) * 2;
          for (int c = 0; c < 2 * nch; c++) acc[2 * (size_t)ri + c] += src[c];
          ri += nch;
        }
      }
      /* copyPCalTones (core.cpp:1124-1157) after Mode::finalisepcal */
      if (!pcx[d].empty()) {
        int64_t pi = lay.pcaloff[d];
        for (int k = 0; k < ds.nrecbands; k++) {
          const int nt = lay.pcal_ntones[d][ds.recband_freq[k]];
          std::vector<cpx<T>> tones(std::max(nt, pcx[d][k].ntones) + 1, cpx<T>{0, 0});
          pcx[d][k].final(tones.data());
          for (int t = 0; t < nt; t++) { acc[2 * (size_t)pi] += tones[t].re; acc[2 * (size_t)pi + 1] += tones[t].im; pi++; }
        }
      }
      int64_t wi = lay.acwoff[d] * 2;
      for (int pass = 0; pass < lay.bandsperautocorr; pass++) {
        if (pass == 1 && !j->crosspol_autocorrs) break;
        for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
          int fi = ds_band_freqindex(ds, k);
          if (!(lay.freqused[fi] || lay.equivused[fi])) continue;
          int parent = k < ds.nrecbands ? k : ds.zoomband_parent[k - ds.nrecbands];
          acc[wi] += w0[d][parent];   /* getWeight(false, ...) for both passes */
          wi++;
        }
      }
    }
  }
  int record_band(int d, int band) const {
    const fxg_datastream& ds = j->datastreams[d];
    return band < ds.nrecbands ? band : ds.zoomband_parent[band - ds.nrecbands];
  }
  const cpx<T>* band_spectrum(int d, int band, const std::vector<std::vector<cpx<T>>>& spec) const {
    const fxg_datastream& ds = j->datastreams[d];
    if (band < ds.nrecbands) return spec[d].data() + (size_t)band * geom[d].C;
    int z = band - ds.nrecbands;
    return spec[d].data() + (size_t)ds.zoomband_parent[z] * geom[d].C + ds.zoomband_choffset[z];
  }
};

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_lut(int format, int nbits, float* out) {
  if (format == FXG_FORMAT_VDIF && nbits == 2) { lut_vdif2(out); return 4; }
  if (format == FXG_FORMAT_VDIF && nbits == 8) { lut_vdif8(out); return 256; }
  if (format == FXG_FORMAT_MARK5B && nbits == 2) { lut_mark5b2(out); return 4; }
  return fail(FXG_EUNSUPPORTED, "lut");
}

int orc_calcstridelength(int n) { return calcstridelength((unsigned)n); }

/* Mk5Mode::unpack of L = nsamp samples at sampleoffset; out[nrecbands][nsamp]. */
int orc_unpack(const fxg_datastream* ds, const uint8_t* data, int64_t nbytes,
               int64_t sampleoffset, int nsamp, float* out, float* weight) {
  StreamFmt f; int r = make_fmt(ds, &f); if (r) return r;
  std::vector<float> work;
  Unpacked u = mk5_unpack(f, data, nbytes, sampleoffset, nsamp, work);
  const int nc = f.cplx ? 2 : 1;     /* complex: (re, im) float pairs */
  for (int c = 0; c < f.nchan; c++)
    std::memcpy(out + (size_t)c * nsamp * nc, work.data() + ((size_t)c * u.stride + u.munged) * nc,
                sizeof(float) * nsamp * nc);
  *weight = u.weight;
  return 0;
}

int orc_plan(const fxg_job_desc* j, fxg_plan_info* info) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  std::memset(info, 0, sizeof(*info));
  info->result_length = L.total;
  info->arraystridelen = L.arraystridelen.empty() ? 0 : L.arraystridelen[0];
  info->xmacstridelen = L.xmacstridelen;
  info->rotatestridelen = L.rotatestridelen;
  info->bandsperautocorr = L.bandsperautocorr;
  info->fftchannels = 2 * j->freqs[j->datastreams[0].recfreq_index[0]].nchan;
  return 0;
}
int orc_plan_baseline_offset(const fxg_job_desc* j, int f, int b, int64_t* off) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (f < 0 || f >= j->nfreqs || b < 0 || b >= j->nbaselines) return fail(FXG_EINVAL, "index");
  *off = L.bloff[f][b]; return 0;
}
int orc_plan_bweight_offset(const fxg_job_desc* j, int f, int b, int64_t* off) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (f < 0 || f >= j->nfreqs || b < 0 || b >= j->nbaselines) return fail(FXG_EINVAL, "index");
  *off = L.bwoff[f][b]; return 0;
}
int orc_plan_autocorr_offset(const fxg_job_desc* j, int d, int64_t* off) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "index");
  *off = L.acoff[d]; return 0;
}
int orc_plan_acweight_offset(const fxg_job_desc* j, int d, int64_t* off) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "index");
  *off = L.acwoff[d]; return 0;
}

int orc_plan_pcal_offset(const fxg_job_desc* j, int d, int64_t* off) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "index");
  *off = L.pcaloff[d]; return 0;
}
int orc_plan_pcal_tones(const fxg_job_desc* j, int d, int recfreq, int32_t* ntones, int32_t* offset_hz) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "index");
  if (L.pcal_ntones[d].empty()) { *ntones = 0; *offset_hz = -1; return 0; }
  if (recfreq < 0 || recfreq >= (int)L.pcal_ntones[d].size()) return fail(FXG_EINVAL, "index");
  *ntones = L.pcal_ntones[d][recfreq]; *offset_hz = L.pcal_offhz[d][recfreq]; return 0;
}

/* One phase-cal extractor over a sample stream, as the reference's unit test
 * drives it (pcal.cpp:1590-1640): kind -1 = PCal::getNew's choice with the
 * given minimum resolution, else PC_*; nseg calls, call i at sample offset
 * so[i] with data + start[i], len[i] samples.  out[*ntones] (complex float
 * pairs); returns the extractor kind. */
int orc_pcal_extract(double bw_hz, double spacing_hz, int offset_hz, int kind, double minres, int nseg,
                     const int64_t* so, const int64_t* start, const int64_t* len, const float* data,
                     float* out, int32_t* ntones, int64_t* samples) {
  if (kind < 0) kind = pcal_kind(bw_hz, spacing_hz, offset_hz, minres);
  PcalExtractor<float> x;
  x.init(kind, bw_hz, spacing_hz, offset_hz, minres);
  for (int i = 0; i < nseg; i++) { x.adjust((size_t)so[i]); x.extract(data + start[i], (size_t)len[i]); }
  std::vector<cpx<float>> o(x.ntones + 1);
  *samples = (int64_t)x.final(o.data());
  *ntones = x.ntones;
  for (int i = 0; i < x.ntones; i++) { out[2 * i] = o[i].re; out[2 * i + 1] = o[i].im; }
  return kind;
}

/* The same for complex samples (data: float pairs, start and len in
 * complex samples): kind -1 = PCal::getNew's choice for complex data, else
 * PC_CSHIFT / PC_CIMPLICIT; dsb = double sideband. */
int orc_pcal_extract_c(double bw_hz, double spacing_hz, int offset_hz, int kind, double minres, int dsb, int nseg,
                       const int64_t* so, const int64_t* start, const int64_t* len, const float* data,
                       float* out, int32_t* ntones, int64_t* samples) {
  if (kind < 0) kind = pcal_kind_c(bw_hz, spacing_hz, offset_hz, minres);
  if (kind != PC_DUMMY && kind != PC_CSHIFT && kind != PC_CIMPLICIT) return fail(FXG_EINVAL, "kind");
  PcalExtractor<float> x;
  x.init(kind, bw_hz, spacing_hz, offset_hz, minres, dsb != 0);
  for (int i = 0; i < nseg; i++) { x.adjust((size_t)so[i]); x.extract(data + 2 * start[i], (size_t)len[i]); }
  std::vector<cpx<float>> o(x.ntones + 1);
  *samples = (int64_t)x.final(o.data());
  *ntones = x.ntones;
  for (int i = 0; i < x.ntones; i++) { out[2 * i] = o[i].re; out[2 * i + 1] = o[i].im; }
  return kind;
}

/* PCal::extractAndIntegrate_reference on one call of len samples. */
int orc_pcal_direct(double bw_hz, double spacing_hz, int offset_hz, int64_t sampleoffset, const float* data,
                    int64_t len, double* out, int32_t* ntones) {
  const int nt = calc_num_tones(bw_hz, (double)offset_hz, spacing_hz) - (offset_hz == 0 ? 1 : 0);
  std::vector<cpx<double>> o(nt + 1);
  uint64_t count = 0;
  pcal_direct(2 * bw_hz, offset_hz, nt, &count, data, (size_t)len, (uint64_t)sampleoffset, o.data());
  for (int i = 0; i < nt; i++) { out[2 * i] = o[i].re; out[2 * i + 1] = o[i].im; }
  *ntones = nt;
  return 0;
}

/* Block geometry of one station/block (the Mode::process prologue). */
int orc_block_info(const fxg_job_desc* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                   int d, int block, fxg_block_info* out) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "datastream");
  DsGeom g; r = make_geom(j, L, d, &g); if (r) return r;
  std::vector<float> work; Unpacked up{};
  BlockGeom bg = block_geom(g, *sub, st[d], block, &work, &up);
  std::memset(out, 0, sizeof(*out));
  out->status = bg.status; out->nearestsample = bg.nearestsample; out->integerdelay = bg.integerdelay;
  out->weight = bg.status == 0 ? bg.weight : 0.0f; out->fracsampleerror = bg.fracsampleerror;
  out->intwalltime = bg.intwalltime; out->a = bg.a; out->b = bg.b; out->fracwalltime = bg.fracwalltime;
  out->c = bg.c;
  return 0;
}

/* Spectra of all recorded bands of station d for one block:
 * out[nrecbands][C] complex, float (fp64 == 0) or double (fp64 != 0). */
int orc_spectra(const fxg_job_desc* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                int d, int block, int fp64, void* out) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "datastream");
  DsGeom g; r = make_geom(j, L, d, &g); if (r) return r;
  if (fp64) {
    FFT<double> fft(g.L); StationWork<double> w;
    station_block<double>(j, g, d, *sub, st[d], block, fft, w, (cpx<double>*)out);
  } else {
    FFT<float> fft(g.L); StationWork<float> w;
    station_block<float>(j, g, d, *sub, st[d], block, fft, w, (cpx<float>*)out);
  }
  return 0;
}

/* Spectral kurtosis records of datastream d for one subintegration
 * (Configuration::dumpKurtosis): per record t of nthreads, the blocks of
 * Core process thread t (Core::updateconfig, core.cpp:2089-2095), the
 * accumulation s1 += |X|^2, s2 += |X|^4 of every recorded band's FP32
 * spectrum before the fractional-sample rotation (mode.cpp:1260-1277; blocks
 * that Mode::process zeroes return before it, :671-684), then
 * calculateAndAverageKurtosis (:1426-1478) with maxchannels =
 * Configuration::getSTADumpChannels.  out[nthreads][nrecbands][nout] with
 * nout = min(maxchannels, nchan); valid[t] = the value it returns (nonzero
 * s1 in the last recorded band).  A band whose s1 is all zero, which the
 * reference skips (leaving the previous call's values), reads zeros. */
int orc_kurtosis(const fxg_job_desc* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                 int d, int nthreads, int maxchannels, float* out, int* valid) {
  Layout L; int r = make_layout(j, &L); if (r) return r;
  if (d < 0 || d >= j->ndatastreams) return fail(FXG_EINVAL, "datastream");
  if (nthreads < 1 || maxchannels < 1) return fail(FXG_EINVAL, "nthreads / maxchannels");
  DsGeom g; r = make_geom(j, L, d, &g); if (r) return r;
  const int nb = j->datastreams[d].nrecbands, C = g.C, n = j->blockspersend;
  const int nout = maxchannels < C ? maxchannels : C;
  FFT<float> fft(g.L); StationWork<float> w;
  std::vector<cpx<float>> spec((size_t)nb * C), pre((size_t)nb * C);
  std::vector<float> s1((size_t)nb * C), s2((size_t)nb * C), sk(C);
  int startblock = 0, numblocks = 0;
  for (int t = 0; t < nthreads; t++) {
    startblock += numblocks;                                       /* core.cpp:2091-2095 */
    numblocks = n / nthreads + ((t < n % nthreads) ? 1 : 0);
    std::fill(s1.begin(), s1.end(), 0.0f);                         /* zeroKurtosis, mode.cpp:1514-1522 */
    std::fill(s2.begin(), s2.end(), 0.0f);
    for (int blk = startblock; blk < startblock + numblocks; blk++) {
      BlockGeom bg = station_block<float>(j, g, d, *sub, st[d], blk, fft, w, spec.data(), pre.data());
      if (bg.status != 0) continue;
      for (int b = 0; b < nb; b++)
        for (int c = 0; c < C; c++) {
          const cpx<float>& x = pre[(size_t)b * C + c];
          float k = std::sqrt(x.re * x.re + x.im * x.im);          /* vectorMagnitude_cf32 */
          k = k * k;                                               /* vectorSquare_f32_I */
          s1[(size_t)b * C + c] += k;
          k = k * k;
          s2[(size_t)b * C + c] += k;
        }
    }
    bool nonzero = false;
    for (int b = 0; b < nb; b++) {
      float* o = out + ((size_t)t * nb + b) * nout;
      float* a = s1.data() + (size_t)b * C;
      float* q = s2.data() + (size_t)b * C;
      nonzero = false;
      for (int c = 0; c < C; c++) if (a[c] > 0.0f) { nonzero = true; break; }
      if (!nonzero) { std::fill(o, o + nout, 0.0f); continue; }
      for (int c = 0; c < C; c++) a[c] *= a[c];
      for (int c = 0; c < C; c++) q[c] *= (float)numblocks;
      for (int c = 0; c < C; c++) sk[c] = q[c] / a[c];
      for (int c = 0; c < C; c++) sk[c] += (float)-1.0;
      const float corr = (float)((float)numblocks / (numblocks - 1.0));
      for (int c = 0; c < C; c++) sk[c] *= corr;
      for (int c = 0; c < C; c++) sk[c] += (float)-1.0;
      if (maxchannels < C) {
        const int kchanavg = C / maxchannels;
        const float inv = (float)(1.0 / ((float)kchanavg));
        for (int c = 0; c < C; c++) sk[c] *= inv;
        for (int jj = 0; jj < maxchannels; jj++) {
          sk[jj] = sk[jj * kchanavg];
          for (int k = 1; k < kchanavg; k++) sk[jj] += sk[jj * kchanavg + k];
        }
      }
      std::copy(sk.begin(), sk.begin() + nout, o);
    }
    valid[t] = nonzero ? 1 : 0;
  }
  return 0;
}

/* Core::processdata over blocks [block0, block0+nblocks) with nthreads
 * threads, each owning a contiguous block range (core.cpp:2091-2095) and its
 * own accumulators, merged into results in thread order.  results has
 * result_length complex values: float pairs (fp64 == 0) or double pairs. */
int orc_process(const fxg_job_desc* j, const fxg_subint_desc* sub, const fxg_station_input* st,
                int fp64, int nthreads, int block0, int nblocks, void* results) {
  if (nthreads < 1) nthreads = 1;
  if (block0 < 0 || nblocks < 0 || block0 + nblocks > j->blockspersend) return fail(FXG_EINVAL, "block range");
  auto go = [&](auto tag) -> int {
    using T = decltype(tag);
    CoreRun<T> core; core.j = j; int r = core.init(); if (r) return r;
    const size_t n = 2 * (size_t)core.lay.total;
    std::vector<std::vector<T>> accs(nthreads, std::vector<T>(n, (T)0));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
      int s0 = block0 + (int)((int64_t)nblocks * t / nthreads);
      int s1 = block0 + (int)((int64_t)nblocks * (t + 1) / nthreads);
      th.emplace_back([&, s0, s1, t]() { core.run_range(*sub, st, s0, s1, accs[t]); });
    }
    for (auto& x : th) x.join();
    T* out = (T*)results;
    for (size_t i = 0; i < n; i++) out[i] = 0;
    for (int t = 0; t < nthreads; t++) for (size_t i = 0; i < n; i++) out[i] += accs[t][i];
    return 0;
  };
  return fp64 ? go(double()) : go(float());
}

/* Forward unnormalised FFT of L complex values (float or double pairs). */
int orc_fft(int L, int fp64, const void* in, void* out) {
  if (L <= 0 || (L & (L - 1))) return fail(FXG_EINVAL, "L");
  if (fp64) {
    FFT<double> f(L); std::vector<std::complex<double>> a((const std::complex<double>*)in, (const std::complex<double>*)in + L);
    f.run(a.data()); std::memcpy(out, a.data(), sizeof(double) * 2 * L);
  } else {
    FFT<float> f(L); std::vector<std::complex<float>> a((const std::complex<float>*)in, (const std::complex<float>*)in + L);
    f.run(a.data()); std::memcpy(out, a.data(), sizeof(float) * 2 * L);
  }
  return 0;
}

/* Model::calculateDelayInterpolator, order 2 (model.cpp:229-304) for a
 * polynomial delay[polyorder+1] (µs, in seconds from the polynomial start):
 * offsettime/timespan in seconds relative to the polynomial start. */
int orc_delay_interpolator(const double* poly, int polyorder, double offsettime, double timespan,
                           int numincrements, double* coeffs) {
  double tp[16]; double samp[3];
  if (polyorder < 0 || polyorder > 14) return fail(FXG_EINVAL, "polyorder");
  auto eval = [&](double dt) {
    tp[0] = 1.0; for (int i = 0; i < polyorder; i++) tp[i + 1] = tp[i] * dt;
    double o = 0; for (int i = 0; i <= polyorder; i++) o += tp[i] * poly[i]; return o;
  };
  samp[1] = eval(offsettime + timespan / 2.0);
  samp[0] = eval(offsettime);
  samp[2] = eval(offsettime + timespan);
  coeffs[0] = (2.0 * samp[0] - 4.0 * samp[1] + 2.0 * samp[2]) / (numincrements * numincrements);
  coeffs[1] = (-3.0 * samp[0] + 4.0 * samp[1] - samp[2]) / numincrements;
  coeffs[2] = samp[0];
  return 0;
}

}  // extern "C"

/* ------------------------------------------------------------------------ */
/* VDIF thread multiplexing: libraries/vdifio/src/vdifmux.c:84-212         */
/* (configurevdifmux, setvdifmuxinputchannels), :346-410                    */
/* (getLowestFrameNumber), :411-470 (getHighestFrameNumber), :507-1044      */
/* (vdifmux), with the corner turners of cornerturners.c stated by their    */
/* common rule: output sample-time t, channel unit k (nBit bits = one       */
/* thread's channels at t) = thread k's unit t, zero for the padding units  */
/* up to the next power of two.  One fanout, non-legacy headers, real data, */
/* no EDV4 validity propagation.                                             */
/* ------------------------------------------------------------------------ */
namespace orcmux {

static const uint32_t FILL = 0x11223344UL;
static const int BAD_THREAD = 20000;

struct Mux {
  int inputFrameSize, inputFramesPerSecond, bitsPerSample, nThread, nSort, nGap, flags;
  int inputChannelsPerThread, nOutputChan, inputDataSize, outputDataSize, outputFrameSize;
  int frameGranularity;
  uint64_t goodMask;
  int chanIndex[1024];
};

static uint32_t w32(const uint8_t* p, int k) { uint32_t v; std::memcpy(&v, p + 4 * k, 4); return v; }
static void put32(uint8_t* p, int k, uint32_t v) { std::memcpy(p + 4 * k, &v, 4); }
static int hframebytes(const uint8_t* h) { return (int)(w32(h, 2) & 0xFFFFFF) * 8; }
static int hnchan(const uint8_t* h) { return 1 << ((w32(h, 2) >> 24) & 0x1F); }
static int hnbits(const uint8_t* h) { return (int)((w32(h, 3) >> 26) & 0x1F) + 1; }
static int hinvalid(const uint8_t* h) { return (int)(w32(h, 0) >> 31); }
static int hthread(const uint8_t* h) { return (int)((w32(h, 3) >> 16) & 0x3FF); }
static int hseconds(const uint8_t* h) { return (int)(w32(h, 0) & 0x3FFFFFFF); }
static int hframe(const uint8_t* h) { return (int)(w32(h, 1) & 0xFFFFFF); }
static int hepoch(const uint8_t* h) { return (int)((w32(h, 1) >> 24) & 0x3F); }

static unsigned gcdu(unsigned u, unsigned v) {      /* vdifmux.c:50-82 (the same value) */
  while (v) { unsigned t = u % v; u = v; v = t; }
  return u;
}

static int configure(Mux* m, int frameSize, int fps, int bits, int nThread, const int* ids, int nSort, int nGap,
                     int flags, int chansPerThread) {
  if (nThread < 1 || nThread > 31) return -3;
  m->inputFrameSize = frameSize; m->inputFramesPerSecond = fps; m->bitsPerSample = bits;
  m->inputChannelsPerThread = chansPerThread;
  m->frameGranularity = (int)(fps / gcdu((unsigned)fps, 1000000000u));
  m->nSort = nSort; m->nGap = nGap < nSort ? nSort : nGap;
  m->nThread = nThread; m->flags = flags;
  int nt = nThread, n = 1;
  while (n < nt) n *= 2;
  m->nOutputChan = n * chansPerThread;
  m->inputDataSize = frameSize - 32;
  /* configurevdifmux sets the output size before setvdifmuxinputchannels
   * multiplies the channel count, so it counts threads only */
  m->outputDataSize = m->inputDataSize * n;
  m->outputFrameSize = m->outputDataSize + 32;
  for (int i = 0; i < 1024; i++) m->chanIndex[i] = BAD_THREAD;
  for (int i = 0; i < nThread; i++) {
    if (ids[i] < 0 || ids[i] > 1023) return -2;
    m->chanIndex[ids[i]] = i;
  }
  m->goodMask = (1ULL << nThread) - 1;
  return 0;
}

/* skip rules shared by the frame scanners and stage 1: returns the step to
 * the next candidate, 0 if a valid frame of the stream starts at i */
static int skip_step(const Mux& m, const uint8_t* cur, int* kind) {
  if (w32(cur, m.inputFrameSize / 4 - 1) == FILL) { *kind = 1; return m.inputFrameSize; }
  if (w32(cur, 0) == FILL) { *kind = 2; return 8; }
  if (hframebytes(cur) != m.inputFrameSize || hnchan(cur) != m.inputChannelsPerThread || hnbits(cur) != m.bitsPerSample) {
    *kind = 3; return 4;
  }
  if ((m.flags & 0x04) && hinvalid(cur) > 0) { *kind = 4; return m.inputFrameSize; }
  *kind = 0;
  return 0;
}

static int64_t frame_number(const Mux& m, const uint8_t* h) {
  return (int64_t)hseconds(h) * m.inputFramesPerSecond + hframe(h);
}

static int64_t lowest_frame(const uint8_t* src, int srcSize, const Mux& m) {
  int64_t lowest = -1; int nFrame = 0;
  for (int i = 0; i <= srcSize - m.inputFrameSize;) {
    int kind, st = skip_step(m, src + i, &kind);
    if (st) { i += st; continue; }
    int64_t fn = frame_number(m, src + i);
    if (lowest < 0 || fn < lowest) lowest = fn;
    ++nFrame; i += m.inputFrameSize;
    if (nFrame >= m.nSort) break;
  }
  return lowest;
}

static int64_t highest_frame(const uint8_t* src, int srcSize, const Mux& m) {
  int64_t highest = -1;
  for (int i = 0; i <= srcSize - m.inputFrameSize;) {
    int kind, st = skip_step(m, src + i, &kind);
    if (st) { i += st; continue; }
    int64_t fn = frame_number(m, src + i);
    if (highest < 0 || fn > highest) highest = fn;
    i += m.inputFrameSize;
  }
  return highest;
}

/* the corner-turn rule: output unit (t, k) of nBit bits */
static void cornerturn(const Mux& m, uint8_t* out, const std::vector<const uint8_t*>& thr) {
  const int nBit = m.bitsPerSample * m.inputChannelsPerThread;
  const int nUnits = m.nOutputChan / m.inputChannelsPerThread;
  const long long totalbits = 8LL * m.outputDataSize;
  std::memset(out, 0, m.outputDataSize);
  for (long long ob = 0; ob < totalbits; ob += nBit) {
    const long long u = ob / nBit;
    const long long t = u / nUnits;
    const int k = (int)(u % nUnits);
    if (k >= m.nThread || thr[k] == nullptr) continue;
    for (int b = 0; b < nBit; b++) {
      const long long ib = t * nBit + b;
      const int bit = (thr[k][ib >> 3] >> (ib & 7)) & 1;
      out[(ob + b) >> 3] |= (uint8_t)(bit << ((ob + b) & 7));
    }
  }
}

/* stats: [0] nValidFrame [1] nInvalidFrame [2] nDiscardedFrame [3] nWrongThread [4] nSkippedByte
 * [5] nFillByte [6] nDuplicateFrame [7] bytesProcessed [8] nGoodFrame [9] nPartialFrame
 * [10] nFillerFrame [11] srcUsed [12] destUsed [13] outputFrameSize [14] nOutputFrame
 * [15] startFrameNumber [16] outputFrameGranularity [17] epoch
 * An invalid output frame's payload is zeroed here; the reference leaves its
 * scratch pointer table there (a decoder skips invalid frames either way). */
static int mux(uint8_t* dest, int destSize, const uint8_t* src, int srcSize, const Mux& m, int64_t startOut,
               int64_t* st) {
  int nValidFrame = 0, nSkip = 0, nFill = 0, nDup = 0, nInvalidFrame = 0, nDiscard = 0, nSeqUnproc = 0;
  int highestDestIndex = 0, bytesProcessed = -1, firstGoodByte = -1, nEnd = 0;
  int nGoodOutput = 0, nBadOutput = 0, nPartialOutput = 0, nWrongThread = 0;
  uint8_t outHeader[32]; std::memset(outHeader, 0, 32);
  int epoch = -1, highestSortedDestIndex = -1, vhUnset = 1, framesIn = 0;
  int nSort = m.nSort;
  const int N = srcSize - m.inputFrameSize;
  if (3 * srcSize <= 4 * nSort * m.inputFrameSize) nSort = (3 * srcSize) / (4 * m.inputFrameSize);
  const int maxSrcIndex = (m.flags & 0x01) ? srcSize - m.inputFrameSize : srcSize - nSort * m.inputFrameSize;
  if ((m.flags & 0x02) && startOut > 0) {
    int64_t r = startOut % m.frameGranularity;
    if (r != 0) startOut += m.frameGranularity - r;
  }
  const int maxDestIndex = destSize / m.outputFrameSize - 1;
  int64_t startFrameNumber = startOut;
  if (startFrameNumber < 0) {
    startFrameNumber = lowest_frame(src, srcSize, m);
    if (startFrameNumber < 0) return -3;
  }
  std::vector<uint64_t> mask(std::max(0, maxDestIndex + 1), 0);
  std::vector<std::vector<const uint8_t*>> ptr(std::max(0, maxDestIndex + 1), std::vector<const uint8_t*>(m.nThread, nullptr));
  int i;
  for (i = 0; i <= N;) {
    const uint8_t* cur = src + i;
    int kind, stp = skip_step(m, cur, &kind);
    if (stp) {
      if (kind == 1) nFill += m.inputFrameSize;
      else if (kind == 2) nFill += 8;
      else if (kind == 3) nSkip += 4;
      else ++nInvalidFrame;
      i += stp;
      continue;
    }
    if (firstGoodByte < 0) firstGoodByte = i;
    const int chanId = m.chanIndex[hthread(cur)];
    if (chanId == BAD_THREAD) { i += m.inputFrameSize; ++nWrongThread; continue; }
    const int64_t frameNumber = frame_number(m, cur);
    if (vhUnset) {
      std::memcpy(outHeader, cur, 16);
      std::memset(outHeader + 16, 0, 16);
      if (m.flags & 0x40) {          /* VDIF_MUX_FLAG_PROPAGATEVALIDITY: EDV4 header (vdifmux.c:666-675) */
        put32(outHeader, 4, (4u << 24) | ((uint32_t)(m.nThread & 0xFF) << 16));
        put32(outHeader, 5, 0xACABFEEDu);
      }
      uint32_t w2 = w32(outHeader, 2), w3 = w32(outHeader, 3), w0 = w32(outHeader, 0);
      unsigned logc = 0; for (int c = m.nOutputChan; c > 1; c /= 2) logc++;
      w2 = (w2 & ~(0x1Fu << 24)) | ((logc & 0x1F) << 24);
      w3 = (w3 & ~(0x3FFu << 16));
      w2 = (w2 & ~0xFFFFFFu) | ((uint32_t)(m.outputFrameSize / 8) & 0xFFFFFF);
      w0 &= 0x7FFFFFFFu;
      put32(outHeader, 0, w0); put32(outHeader, 2, w2); put32(outHeader, 3, w3);
      epoch = hepoch(outHeader);
      vhUnset = 0;
    }
    ++framesIn;
    const int64_t destIndex64 = frameNumber - startFrameNumber;
    const int destIndex = (int)destIndex64;
    if (destIndex64 < 0) { i += m.inputFrameSize; ++nDiscard; continue; }
    if (i > maxSrcIndex && highestSortedDestIndex < 0) highestSortedDestIndex = highestDestIndex;
    if (destIndex64 > maxDestIndex) {
      ++nSeqUnproc;
      if (framesIn < nSort) { i += m.inputFrameSize; ++nDiscard; continue; }
      if (bytesProcessed == -1) bytesProcessed = i;
      i += m.inputFrameSize;
      ++nEnd;
      if (nEnd >= nSort) break;
    } else {
      --nSeqUnproc;
      if (destIndex > highestDestIndex + m.nGap && m.nGap > 0) {
        if (nValidFrame > nSort) {
          if (bytesProcessed == -1) bytesProcessed = i;
          ++nEnd;
          if (nEnd >= nSort) break;
        }
      }
      if (mask[destIndex] & (1ULL << chanId)) ++nDup;
      else {
        mask[destIndex] |= 1ULL << chanId;
        ptr[destIndex][chanId] = cur + 32;
        ++nValidFrame;
        if (destIndex > highestDestIndex) highestDestIndex = destIndex;
      }
      i += m.inputFrameSize;
    }
  }
  if (bytesProcessed == -1) bytesProcessed = i;
  const bool propagate = (m.flags & 0x40) != 0;
  auto backoff = [&](int f) {
    if (propagate) {                 /* only frames with no thread (vdifmux.c:799-805, 841-847) */
      if (mask[f] == 0) highestDestIndex = f - 1;
      return;
    }
    if (mask[f] != m.goodMask) {
      highestDestIndex = f - 1;
      for (int t = 0; t < m.nThread; ++t)
        if (mask[f] & (1ULL << t)) {
          int d = (int)(ptr[f][t] - src - 32);
          if (d < bytesProcessed) bytesProcessed = d;
          --nValidFrame;
        }
    }
  };
  if (highestSortedDestIndex >= 0) {
    for (int f = highestDestIndex; f > highestSortedDestIndex; --f) backoff(f);
  } else {
    const int minDestIndex = highestDestIndex - nSort / m.nThread;
    if (minDestIndex >= 0) for (int f = highestDestIndex; f >= minDestIndex; --f) backoff(f);
  }
  if (i < N && highestDestIndex < maxDestIndex && (highestDestIndex > maxDestIndex - m.nGap || m.nGap == 0))
    highestDestIndex = maxDestIndex;
  int seconds = (int)(startFrameNumber / m.inputFramesPerSecond);
  int frameNum = (int)(startFrameNumber % m.inputFramesPerSecond);
  for (int f = 0; f <= highestDestIndex; ++f) {
    uint8_t* frame = dest + (size_t)m.outputFrameSize * f;
    std::memcpy(frame, outHeader, 32);
    put32(frame, 0, (w32(frame, 0) & ~0x3FFFFFFFu) | ((uint32_t)seconds & 0x3FFFFFFF));
    put32(frame, 1, (w32(frame, 1) & ~0xFFFFFFu) | ((uint32_t)frameNum & 0xFFFFFF));
    if (propagate && mask[f] != 0) {
      /* the validity mask in header words 6-7; a missing thread's units are
       * zero here (the reference points them at the start of the input,
       * vdifmux.c:929-937) */
      put32(frame, 6, (uint32_t)(mask[f] & 0xFFFFFFFFull));
      put32(frame, 7, (uint32_t)(mask[f] >> 32));
      cornerturn(m, frame + 32, ptr[f]);
      if (mask[f] == m.goodMask) ++nGoodOutput; else ++nPartialOutput;
    } else if (mask[f] == m.goodMask) {
      cornerturn(m, frame + 32, ptr[f]);
      ++nGoodOutput;
    } else {
      put32(frame, 0, w32(frame, 0) | 0x80000000u);
      std::memset(frame + 32, 0, m.outputDataSize);
      ++nBadOutput;
    }
    if (++frameNum >= m.inputFramesPerSecond) { ++seconds; frameNum -= m.inputFramesPerSecond; }
  }
  int64_t srcUsed = bytesProcessed, nOut = nGoodOutput + nBadOutput + nPartialOutput, startFN = startFrameNumber;
  if (nOut == 0 || bytesProcessed == 0 || firstGoodByte < 0 || nSeqUnproc >= nSort) {
    int64_t h = highest_frame(src, srcSize, m);
    startFN = h < 0 ? startFN + (srcSize / m.inputFrameSize) / m.nThread + 1 : h + 1;
    nOut = 0;
    srcUsed = srcSize - (srcSize % m.inputFrameSize);
    bytesProcessed = (int)srcUsed;
  }
  const int64_t vals[18] = {nValidFrame, nInvalidFrame, nDiscard, nWrongThread, nSkip, nFill, nDup,
                            bytesProcessed, nGoodOutput, nPartialOutput, nBadOutput, srcUsed,
                            (int64_t)(nGoodOutput + nBadOutput + nPartialOutput) * m.outputFrameSize,
                            m.outputFrameSize, nOut, startFN, m.frameGranularity, epoch};
  for (int k = 0; k < 18; k++) st[k] = vals[k];
  return bytesProcessed > firstGoodByte ? bytesProcessed : firstGoodByte;
}

}  // namespace orcmux

extern "C" int orc_vdifmux(uint8_t* dest, int destSize, const uint8_t* src, int srcSize, int inputFrameSize,
                           int framesPerSecond, int bitsPerSample, int nThread, const int* threadIds, int nSort,
                           int nGap, int flags, int channelsPerThread, int64_t startOutputFrame, int64_t* stats) {
  if (flags & ~0x47) return -4;                      /* legacy and complex multiplexing: not restated */
  orcmux::Mux m;
  int r = orcmux::configure(&m, inputFrameSize, framesPerSecond, bitsPerSample, nThread, threadIds, nSort, nGap, flags,
                            channelsPerThread);
  if (r) return r;
  return orcmux::mux(dest, destSize, src, srcSize, m, startOutputFrame, stats);
}
