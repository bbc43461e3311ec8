// datastream.cpp — the station data boundary (include/fxcorr_datastream.h):
// DataStream::calculateControlParams (mpifxcorr/src/datastream.cpp:376-630)
// followed by Mk5DataStream's move to a frame boundary (mk5.cpp:146-268),
// for a stream held whole in one device buffer.  The reference's ring keeps
// readbytes-long segments and walks them; here the one segment is the whole
// buffer, so its "next segment" is itself and the continuity clauses that
// compare neighbouring segments never apply.
#include "fxcorr_datastream.h"

#include <algorithm>
#include <cmath>
#include <cstring>

struct fxds_stream {
  fxds_format f{};
  const uint8_t* data = nullptr;
  long long bytes = 0;           // bufferbytes = readbytes = validbytes of the one segment
  int scan = 0, sec = 0, ns = 0; // bufferinfo scan / scanseconds / scanns
  double sampletimens = 0;
  int bytesbetweenintegerns = 0;
  long long nsinc = 0;
};

namespace {

constexpr int kFlagsPerInt = FXG_FLAGS_PER_INT;
constexpr double kTiny = 0.000000001;   // mode.cpp:35

long long gcd_ll(long long a, long long b) { while (b) { long long t = a % b; a = b; b = t; } return a; }

int recfreq_of(const fxg_job_desc* j, int d, const fxg_freq** f) {
  if (!j || !j->datastreams || !j->freqs || d < 0 || d >= j->ndatastreams) return FXG_EINVAL;
  const fxg_datastream& ds = j->datastreams[d];
  if (ds.nrecfreqs < 1 || ds.recfreq_index[0] < 0 || ds.recfreq_index[0] >= j->nfreqs) return FXG_EINVAL;
  *f = &j->freqs[ds.recfreq_index[0]];
  return 0;
}

}  // namespace

extern "C" {

int64_t fxds_data_bytes(const fxg_job_desc* j, int d, double guard_ns) {
  const fxg_freq* f = nullptr;
  if (recfreq_of(j, d, &f)) return FXG_EINVAL;
  const fxg_datastream& ds = j->datastreams[d];
  // decimation factor 1 (configuration.cpp:807-842)
  long long valid = ((long long)j->blockspersend * ds.nrecbands * 2 * ds.nbits * f->nchan) / 8;
  valid += ((long long)(int)(guard_ns / (1000.0 / (f->bw_mhz * 2.0)) + 1.0) * ds.nrecbands * ds.nbits) / 8;
  const long long payload = ds.framebytes - ds.headerbytes;
  long long nframes = valid / payload + 2;
  if (ds.format == FXG_FORMAT_MARK5B) nframes += nframes % 2;
  return nframes * ds.framebytes;
}

int fxds_format_of(const fxg_job_desc* j, int d, double guard_ns, fxds_format* out) {
  const fxg_freq* f = nullptr;
  if (!out || recfreq_of(j, d, &f)) return FXG_EINVAL;
  const fxg_datastream& ds = j->datastreams[d];
  fxds_format r{};
  r.framebytes = ds.framebytes;
  r.payloadbytes = ds.framebytes - ds.headerbytes;                      // getFramePayloadBytes
  const double samplerate = 2.0e6 * f->bw_mhz;
  r.framespersecond = samplerate * ds.nrecbands * ds.nbits / (8.0 * r.payloadbytes);   // getFramesPerSecond
  const long long fps = (long long)std::llround(r.framespersecond);
  r.framegranularity = fps > 0 ? (int)(fps / gcd_ll(fps, 1000000000LL)) : 1;
  r.bytespersamplenum = (ds.nrecbands * ds.nbits) / 8;                   // configuration.cpp:1695-1702
  r.bytespersampledenom = 1;
  if (r.bytespersamplenum == 0) { r.bytespersamplenum = 1; r.bytespersampledenom = 8 / (ds.nrecbands * ds.nbits); }
  r.numchannels = f->nchan;
  r.blockspersend = j->blockspersend;
  r.bw_mhz = f->bw_mhz;
  r.databytes = fxds_data_bytes(j, d, guard_ns);
  *out = r;
  return 0;
}

int fxds_create(const fxds_format* f, fxds_stream** out) {
  if (!f || !out || f->framebytes <= 0 || f->payloadbytes <= 0 || f->framespersecond <= 0 || f->bw_mhz <= 0 ||
      f->bytespersamplenum <= 0 || f->bytespersampledenom <= 0 || f->numchannels <= 0 || f->blockspersend <= 0)
    return FXG_EINVAL;
  fxds_stream* s = new fxds_stream();
  s->f = *f;
  if (s->f.framegranularity < 1) s->f.framegranularity = 1;
  s->sampletimens = 500.0 / f->bw_mhz;                                 // datastream.cpp:763
  double acc = 0.0;                                                     // :765-770
  do {
    acc += f->bytespersampledenom * s->sampletimens;
    s->bytesbetweenintegerns += f->bytespersamplenum;
  } while (!(std::fabs(acc - int(acc + 0.5)) < kTiny));
  *out = s;
  return 0;
}

int fxds_destroy(fxds_stream* s) { delete s; return 0; }

int fxds_set_buffer(fxds_stream* s, const uint8_t* p, int64_t bytes, int32_t scan, int32_t sec, int32_t ns) {
  if (!s || (!p && bytes > 0) || bytes < 0) return FXG_EINVAL;
  s->data = p; s->bytes = bytes; s->scan = scan; s->sec = sec; s->ns = ns;
  // Mk5DataStream::updateConfig (mk5.cpp:284): whole frames times the frame time
  s->nsinc = (long long)(int)((double)(bytes / s->f.framebytes) * (1000000000.0 / s->f.framespersecond) + 0.5);
  return 0;
}

int fxds_control(fxds_stream* s, int32_t scan, int32_t offsetsec, int32_t offsetns, double delayus1, double delayus2,
                 int32_t model_ok, fxg_station_input* out, int32_t* flags, int32_t nflags, int64_t* bufindex) {
  if (!s || !out || !flags) return FXG_EINVAL;
  const fxds_format& F = s->f;
  const int nwords = (F.blockspersend + kFlagsPerInt - 1) / kFlagsPerInt;
  if (nflags < nwords) return FXG_EINVAL;
  std::memset(out, 0, sizeof(*out));
  for (int i = 0; i < nflags; i++) flags[i] = 0;
  out->scan = scan;
  out->sec = FXG_INVALID_SUBINT;
  out->nvalidflags = nwords;
  out->validflags = flags;
  if (bufindex) *bufindex = 0;
  auto torch = [&]() { out->sec = FXG_INVALID_SUBINT; out->data = nullptr; out->nbytes = 0; return 0; };
  if (!model_ok) return torch();

  // ---- DataStream::calculateControlParams (datastream.cpp:376-630), one segment
  const long long readbytes = s->bytes, bufferbytes = s->bytes, validbytes = s->bytes;
  long long firstoffsetns = (long long)offsetns - static_cast<long long>(delayus1 * 1000);
  const long long dataspanns = static_cast<long long>(F.numchannels * F.blockspersend * 2 * s->sampletimens + 0.5);
  long long lastoffsetns = (long long)offsetns + dataspanns + static_cast<long long>(-1000 * delayus2 + 0.5);
  for (int k = 0; k < 2; k++)
    if (lastoffsetns < 0) { offsetsec -= 1; firstoffsetns += 1000000000LL; lastoffsetns += 1000000000LL; }
  if (scan < s->scan) return torch();
  if (readbytes <= 0) return torch();                                 // nothing ever read
  if (scan == s->scan && offsetsec < s->sec - 3) return torch();
  const long long nsdifference = (long long)(offsetsec - s->sec) * 1000000000LL + firstoffsetns - (long long)s->ns;
  const long long validns = (validbytes * s->nsinc) / readbytes;
  // past the end of the one segment: no later segment to move to
  if (scan > s->scan || (scan == s->scan && nsdifference >= validns)) return torch();
  if (nsdifference < -s->nsinc) return torch();
  const int fftchannels = F.numchannels * 2;
  const int blockbytes = (fftchannels * F.bytespersamplenum) / F.bytespersampledenom;
  // (the reference holds this in an int: its buffers stay below 2 GB)
  long long bufferindex =
      (static_cast<long long>(((long long)(offsetsec - s->sec) * 1000000000LL + (firstoffsetns - (long long)s->ns)) /
                                  s->sampletimens + 0.5) * F.bytespersamplenum) / F.bytespersampledenom;
  if (bufferindex % 2 != 0) bufferindex--;                            // previous 16-bit boundary
  if (bufferindex == bufferbytes) return torch();                     // at the end, and no continuation
  int count = 0;
  while (bufferindex < 0 && count < F.blockspersend) { bufferindex += blockbytes; count++; }
  if (bufferindex < 0) return torch();
  if (count > 0) {
    // skipped blocks: allow for the delay changing over the subint (:545-560)
    int tosubtract = (int(count * 1000.0 * (delayus2 - delayus1) / (s->sampletimens * F.blockspersend) + 0.5) *
                      F.bytespersamplenum) / F.bytespersampledenom;
    if (bufferindex - tosubtract < 0) {
      count++;
      bufferindex += blockbytes;
      if (count == F.blockspersend) return torch();
      tosubtract = (int(count * 1000.0 * (delayus2 - delayus1) / (s->sampletimens * F.blockspersend) + 0.5) *
                    F.bytespersamplenum) / F.bytespersampledenom;
      bufferindex -= tosubtract;
    }
    bufferindex -= bufferindex % 2;
  }
  long long segoffbytes = bufferindex;
  bufferindex -= segoffbytes % s->bytesbetweenintegerns;
  segoffbytes -= segoffbytes % s->bytesbetweenintegerns;
  if (validbytes - segoffbytes >= F.databytes) {
    for (int i = count; i < F.blockspersend; i++) flags[i / kFlagsPerInt] |= 1 << (i % kFlagsPerInt);
  } else {
    const double dbufferindex = ((((double)(offsetsec - s->sec)) * 1.0e9 + (double)(firstoffsetns - (long long)s->ns)) /
                                 s->sampletimens) * F.bytespersamplenum / F.bytespersampledenom;
    if (dbufferindex < -(double)blockbytes * F.blockspersend || dbufferindex > (double)bufferbytes) return torch();
    for (int i = count; i < F.blockspersend; i++)
      if (bufferindex + (long long)i * blockbytes < validbytes) flags[i / kFlagsPerInt] |= 1 << (i % kFlagsPerInt);
    bool somegood = false;                    // full flag words only, as the reference checks
    for (int i = 0; i < F.blockspersend / kFlagsPerInt; i++)
      if (flags[i] > 0) somegood = true;
    if (!somegood) return torch();
  }

  // ---- Mk5DataStream: back to a frame boundary on the frame granularity (mk5.cpp:208-262)
  long long framesin = bufferindex / F.payloadbytes;
  if (framesin % F.framegranularity != 0) {
    framesin -= framesin % F.framegranularity;
    if (framesin < 0) framesin += F.framegranularity;
  }
  const long long segoffns = (long long)s->ns + (long long)((1000000000.0 * framesin) / F.framespersecond + 0.5);
  const int datasec = s->sec + (int)(segoffns / 1000000000LL);
  const int datans = (int)(segoffns % 1000000000LL);
  bufferindex = framesin * F.framebytes;
  if (bufferindex >= bufferbytes) return torch();
  out->sec = datasec;
  out->ns = datans;
  out->data = s->data + bufferindex;
  out->nbytes = std::min<long long>(F.databytes, bufferbytes - bufferindex);
  if (bufindex) *bufindex = bufferindex;
  return 0;
}

}  // extern "C"
