// fx_mux.cpp — VDIF thread multiplexing, host side: configuration
// (configurevdifmux / setvdifmuxinputchannels, libraries/vdifio/src/
// vdifmux.c:84-253) and the frame search and ordering of vdifmux (stage 1,
// :507-876), which reads only the 32-byte headers.  The result is a table
// of output frames (header, threads present, each thread's payload offset)
// that the corner-turn kernel (fx_mux.hip) turns into device frames.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fx_kernels.h"

namespace fxg {

namespace {

constexpr uint32_t kFill = 0x11223344u;     // fill pattern, little endian (vdifmux.c:40-44)

struct MuxConfig {
  int inframe = 0, indata = 0, fps = 0, bits = 0, nthreads = 0, chans = 1;
  int nsort = 0, ngap = 0, flags = 0;
  int units = 1;                 // thread units per sample-time: next power of two >= nthreads
  int outchans = 1, outdata = 0, outframe = 0, granularity = 1;
  unsigned long long good = 0;
  std::vector<int> chan_of_thread;   // VDIF thread id -> input index, -1 if not wanted
};

inline uint32_t word(const uint8_t* p, int k) { uint32_t v; std::memcpy(&v, p + 4 * k, 4); return v; }
inline void setword(uint8_t* p, int k, uint32_t v) { std::memcpy(p + 4 * k, &v, 4); }

long long gcd_ll(long long a, long long b) { while (b) { long long t = a % b; a = b; b = t; } return a; }

int configure(const fxg_vdifmux_desc* d, MuxConfig* c, std::string* err) {
  if (!d) { *err = "null descriptor"; return FXG_EINVAL; }
  if (d->nthreads < 1 || d->nthreads > FXG_MUX_MAX_THREADS) { *err = "nthreads out of range"; return FXG_EINVAL; }
  if (d->flags & ~(FXG_MUX_GOTOEND | FXG_MUX_RESPECTGRANULARITY | FXG_MUX_ENABLEVALIDITY | FXG_MUX_PROPAGATEVALIDITY)) {
    *err = "legacy and complex multiplexing are not built"; return FXG_EUNSUPPORTED;
  }
  if (d->input_frame_bytes <= 32 || d->input_frame_bytes % 8 || d->frames_per_second <= 0 || d->nsort < 0 ||
      d->channels_per_thread < 1) { *err = "frame size, rate, nsort or channels"; return FXG_EINVAL; }
  const int unitbits = d->bits_per_sample * d->channels_per_thread;
  if (unitbits < 1 || unitbits > 32 || (unitbits & (unitbits - 1))) {
    *err = "bits per sample x channels per thread must be a power of two <= 32"; return FXG_EUNSUPPORTED;
  }
  c->inframe = d->input_frame_bytes; c->indata = c->inframe - 32;
  c->fps = d->frames_per_second; c->bits = d->bits_per_sample; c->nthreads = d->nthreads;
  c->chans = d->channels_per_thread; c->nsort = d->nsort; c->ngap = std::max(d->ngap, d->nsort);
  c->flags = d->flags;
  c->granularity = (int)(c->fps / gcd_ll(c->fps, 1000000000LL));
  c->units = 1;
  while (c->units < c->nthreads) c->units *= 2;
  c->outchans = c->units * c->chans;
  c->outdata = c->indata * c->units;       // set before the channel factor (vdifmux.c:177, 241)
  c->outframe = c->outdata + 32;
  c->good = (1ULL << c->nthreads) - 1;
  c->chan_of_thread.assign(1024, -1);
  for (int i = 0; i < c->nthreads; i++) {
    if (d->thread_ids[i] < 0 || d->thread_ids[i] > 1023) { *err = "thread id out of range"; return FXG_EINVAL; }
    c->chan_of_thread[d->thread_ids[i]] = i;
  }
  return 0;
}

// A frame of the wanted stream starts at p (step 0), or how far to move on and why.
enum Skip { SK_NONE, SK_FILLEND, SK_FILLSTART, SK_SYNC, SK_INVALID };
int skip(const MuxConfig& c, const uint8_t* p, Skip* why) {
  if (word(p, c.inframe / 4 - 1) == kFill) { *why = SK_FILLEND; return c.inframe; }
  if (word(p, 0) == kFill) { *why = SK_FILLSTART; return 8; }
  const uint32_t w2 = word(p, 2), w3 = word(p, 3);
  if ((int)(w2 & 0xFFFFFF) * 8 != c.inframe || (1 << ((w2 >> 24) & 0x1F)) != c.chans ||
      (int)((w3 >> 26) & 0x1F) + 1 != c.bits) { *why = SK_SYNC; return 4; }
  if ((c.flags & FXG_MUX_ENABLEVALIDITY) && (word(p, 0) >> 31)) { *why = SK_INVALID; return c.inframe; }
  *why = SK_NONE;
  return 0;
}

long long frame_no(const MuxConfig& c, const uint8_t* p) {
  return (long long)(word(p, 0) & 0x3FFFFFFF) * c.fps + (word(p, 1) & 0xFFFFFF);
}

// lowest (first nsort frames) or highest (all frames) frame number, -1 if none
long long scan_frames(const MuxConfig& c, const uint8_t* src, long long n, bool lowest) {
  long long best = -1;
  int seen = 0;
  for (long long i = 0; i <= n - c.inframe;) {
    Skip why;
    int st = skip(c, src + i, &why);
    if (st) { i += st; continue; }
    const long long fn = frame_no(c, src + i);
    if (best < 0 || (lowest ? fn < best : fn > best)) best = fn;
    i += c.inframe;
    if (lowest && ++seen >= c.nsort) break;
  }
  return best;
}

}  // namespace

}  // namespace fxg

using namespace fxg;

extern "C" {

int fxg_vdifmux_output_shape(const fxg_vdifmux_desc* d, int32_t* frame_bytes, int32_t* nchan) {
  MuxConfig c;
  std::string err;
  int rc = configure(d, &c, &err);
  if (rc) return set_error(rc, err.c_str());
  if (frame_bytes) *frame_bytes = c.outframe;
  if (nchan) *nchan = c.outchans;
  return 0;
}

int64_t fxg_vdifmux(fxg_ctx* ctx, const fxg_vdifmux_desc* d, const uint8_t* src, const uint8_t* src_device,
                    int64_t srcsize, uint8_t* dest, int64_t destsize, int64_t start, fxg_vdifmux_stats* stats) {
  MuxConfig c;
  std::string err;
  int rc = configure(d, &c, &err);
  if (rc) return set_error(rc, err.c_str());
  if (!ctx || !src || !dest || srcsize < 0 || destsize < 0) return set_error(FXG_EINVAL, "null argument");
  if (((uintptr_t)src_device | (uintptr_t)dest) % 4) return set_error(FXG_EINVAL, "device buffers must be 4-byte aligned");
  long long nvalid = 0, nskip = 0, nfill = 0, ndup = 0, ninvalid = 0, ndiscard = 0, nwrong = 0;
  long long seq_unprocessable = 0, highest = 0, processed = -1, firstgood = -1, nend = 0;
  long long highest_sorted = -1, frames_in = 0;
  bool have_header = false;
  uint8_t proto[32] = {0};
  int epoch = -1;
  const long long N = srcsize - c.inframe;
  long long nsort = c.nsort;
  if (3 * srcsize <= 4 * nsort * c.inframe) nsort = (3 * srcsize) / (4 * c.inframe);
  const long long max_src = (c.flags & FXG_MUX_GOTOEND) ? srcsize - c.inframe : srcsize - nsort * c.inframe;
  if ((c.flags & FXG_MUX_RESPECTGRANULARITY) && start > 0 && start % c.granularity)
    start += c.granularity - start % c.granularity;
  const long long max_dest = destsize / c.outframe - 1;
  long long first = start;
  if (first < 0) {
    first = scan_frames(c, src, srcsize, true);
    if (first < 0) return set_error(FXG_EINVAL, "no frame of the stream in the input");
  }
  const long long nout = std::max(0LL, max_dest + 1);
  std::vector<unsigned long long> mask(nout, 0);
  std::vector<long long> off((size_t)nout * c.nthreads, 0);

  // ---- stage 1: place each frame of the wanted threads on the output grid
  long long i = 0;
  while (i <= N) {
    const uint8_t* p = src + i;
    Skip why;
    int st = skip(c, p, &why);
    if (st) {
      if (why == SK_FILLEND) nfill += c.inframe;
      else if (why == SK_FILLSTART) nfill += 8;
      else if (why == SK_SYNC) nskip += 4;
      else ninvalid++;
      i += st;
      continue;
    }
    if (firstgood < 0) firstgood = i;
    const int ch = c.chan_of_thread[(word(p, 3) >> 16) & 0x3FF];
    if (ch < 0) { i += c.inframe; nwrong++; continue; }
    if (!have_header) {
      // the output header: the first frame's, 16 extended bytes cleared (or,
      // propagating validity, EDV4: mask length = threads, version 4, sync
      // word, the mask per frame below), the channel count, thread 0, the
      // output length, valid (vdifmux.c:664-690)
      std::memcpy(proto, p, 16);
      if (c.flags & FXG_MUX_PROPAGATEVALIDITY) {
        setword(proto, 4, (4u << 24) | ((uint32_t)(c.nthreads & 0xFF) << 16));
        setword(proto, 5, 0xACABFEEDu);
      }
      unsigned logc = 0;
      for (int n = c.outchans; n > 1; n /= 2) logc++;
      setword(proto, 0, word(proto, 0) & 0x7FFFFFFFu);
      setword(proto, 2, (word(proto, 2) & 0xE0000000u) | ((logc & 0x1F) << 24) | ((uint32_t)(c.outframe / 8) & 0xFFFFFF));
      setword(proto, 3, word(proto, 3) & ~(0x3FFu << 16));
      epoch = (int)((word(proto, 1) >> 24) & 0x3F);
      have_header = true;
    }
    frames_in++;
    const long long di = frame_no(c, p) - first;
    if (di < 0) { i += c.inframe; ndiscard++; continue; }
    if (i > max_src && highest_sorted < 0) highest_sorted = highest;
    if (di > max_dest) {
      seq_unprocessable++;
      if (frames_in < nsort) { i += c.inframe; ndiscard++; continue; }
      if (processed < 0) processed = i;
      i += c.inframe;
      if (++nend >= nsort) break;
      continue;
    }
    seq_unprocessable--;
    if (di > highest + c.ngap && c.ngap > 0 && nvalid > nsort) {
      if (processed < 0) processed = i;
      if (++nend >= nsort) break;
    }
    if (mask[di] & (1ULL << ch)) {
      ndup++;
    } else {
      mask[di] |= 1ULL << ch;
      off[di * c.nthreads + ch] = i + 32;
      nvalid++;
      highest = std::max(highest, di);
    }
    i += c.inframe;
  }
  if (processed < 0) processed = i;
  // the source dried up: drop the incomplete frames at the end so that the
  // next call completes them (vdifmux.c:785-873)
  const bool propagate = (c.flags & FXG_MUX_PROPAGATEVALIDITY) != 0;
  auto drop_incomplete = [&](long long f) {
    if (propagate) {                     // only frames with no thread at all (vdifmux.c:799-805)
      if (mask[f] == 0) highest = f - 1;
      return;
    }
    if (mask[f] == c.good) return;
    highest = f - 1;
    for (int t = 0; t < c.nthreads; t++)
      if (mask[f] & (1ULL << t)) {
        processed = std::min(processed, off[f * c.nthreads + t] - 32);
        nvalid--;
      }
  };
  if (highest_sorted >= 0) {
    for (long long f = highest; f > highest_sorted; --f) drop_incomplete(f);
  } else {
    const long long lo = highest - nsort / c.nthreads;
    if (lo >= 0) for (long long f = highest; f >= lo; --f) drop_incomplete(f);
  }
  if (i < N && highest < max_dest && (highest > max_dest - c.ngap || c.ngap == 0)) highest = max_dest;

  // ---- stage 2: headers on the host, payloads on the GPU
  const long long nframes = nout > 0 ? highest + 1 : 0;
  long long good = 0, filler = 0, partial = 0;
  std::vector<uint32_t> hdr((size_t)nframes * 8);
  long long sec = first / c.fps, fno = first % c.fps;
  for (long long f = 0; f < nframes; f++) {
    uint32_t* h = hdr.data() + f * 8;
    std::memcpy(h, proto, 32);
    h[0] = (h[0] & ~0x3FFFFFFFu) | ((uint32_t)sec & 0x3FFFFFFF);
    h[1] = (h[1] & ~0xFFFFFFu) | ((uint32_t)fno & 0xFFFFFF);
    if (propagate && mask[f] != 0) {
      // the threads present, as the EDV4 validity mask (vdifmux.c:899-927)
      h[6] = (uint32_t)(mask[f] & 0xFFFFFFFFull);
      h[7] = (uint32_t)(mask[f] >> 32);
      if (mask[f] == c.good) good++; else partial++;
    } else if (mask[f] == c.good) {
      good++;
    } else {
      h[0] |= 0x80000000u;
      filler++;
    }
    if (++fno >= c.fps) { sec++; fno -= c.fps; }
  }
  if (nframes > 0) {
    rc = ctx_mux_launch(ctx, src, src_device, srcsize, dest, hdr.data(), mask.data(), off.data(), nframes,
                        c.nthreads, c.bits * c.chans, c.units, c.outdata / 4, c.outframe, c.good, propagate ? 1 : 0);
    if (rc) return rc;
  }
  long long src_used = processed, nout_frames = good + partial + filler, start_fn = first;
  if (nout_frames == 0 || processed == 0 || firstgood < 0 || seq_unprocessable >= nsort) {
    const long long h = scan_frames(c, src, srcsize, false);
    start_fn = h < 0 ? start_fn + (srcsize / c.inframe) / c.nthreads + 1 : h + 1;
    nout_frames = 0;
    src_used = srcsize - srcsize % c.inframe;
    processed = src_used;
  }
  if (stats) {
    fxg_vdifmux_stats s{};
    s.valid_frames = nvalid; s.invalid_frames = ninvalid; s.discarded_frames = ndiscard;
    s.wrong_thread_frames = nwrong; s.skipped_bytes = nskip; s.fill_bytes = nfill; s.duplicate_frames = ndup;
    s.bytes_processed = processed; s.good_frames = good; s.partial_frames = partial; s.filler_frames = filler;
    s.src_used = src_used; s.dest_used = (good + partial + filler) * c.outframe; s.output_frame_bytes = c.outframe;
    s.output_frames = nout_frames; s.start_frame_number = start_fn; s.output_frame_granularity = c.granularity;
    s.epoch = epoch;
    *stats = s;
  }
  return processed > firstgood ? processed : firstgood;
}

}  // extern "C"
