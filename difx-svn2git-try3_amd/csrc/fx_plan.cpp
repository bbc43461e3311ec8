// fx_plan.cpp — host-side job planning for libfxcorr_gpu: descriptor checks,
// stride selection, the core result layout and the device-side tables.
//
// Result layout follows Configuration::populateResultLengths
// (mpifxcorr/src/configuration.cpp:2212-2441) for one configuration with a
// single phase centre and no pulsar binning:
//   cross-correlations  [freq][baseline][polproduct][chan]   (cf32)
//   baseline weights    [freq][baseline] ceil(npol/2) cf32 slots of floats
//   autocorrelations    [datastream][pass][total band][chan]  (cf32)
//   autocorr weights    [datastream] ceil(nbands*pass/2) cf32 slots of floats
//   phase cal           [datastream][recorded band][tone]            (cf32)
// Used frequencies are those named by any baseline; a frequency that is not
// used but matches a used one (configuration.cpp:345-367) still gets its
// autocorrelations.
#include "fx_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <map>
#include <array>

namespace fxg {

namespace {

// smallest divisor of n that is >= sqrt(n) (configuration.cpp:46-59)
int stride_for(unsigned n) {
  if (n <= 16) return (int)n;
  unsigned g = (unsigned)std::sqrt(n + 0.1);
  while (n % g) --g;
  return (int)(n / g);
}

// divisor of the channel-count gcd closest to 150 (configuration.cpp:2464-2511)
int xmac_stride_for(const fxg_job_desc* d) {
  long g = 0;
  for (int b = 0; b < d->nbaselines; b++)
    for (int f = 0; f < d->baselines[b].nfreqs; f++) {
      long n = d->freqs[d->baselines[b].freq_index[f]].nchan;
      g = g ? std::gcd(g, n) : n;
    }
  int best = 1;
  double err = 1e12;
  for (long c = g; c > 0; --c) {
    if (g % c) continue;
    double e = std::fabs((double)c - 150.0);
    if (e < err) { err = e; best = (int)c; }
  }
  return best;
}

int freq_of_band(const fxg_datastream& ds, int k) {
  return k < ds.nrecbands ? ds.recfreq_index[ds.recband_freq[k]]
                          : ds.zoomband_freqindex[k - ds.nrecbands];
}

int parent_of_band(const fxg_datastream& ds, int k) {
  return k < ds.nrecbands ? k : ds.zoomband_parent[k - ds.nrecbands];
}

int choff_of_band(const fxg_datastream& ds, int k) {
  return k < ds.nrecbands ? 0 : ds.zoomband_choffset[k - ds.nrecbands];
}

bool power_of_two(long v) { return v > 0 && (v & (v - 1)) == 0; }

// f and g cover the same sky band with opposite sidebands (the USB x LSB
// products of configuration.cpp:2926-2937): equal lower edges and bandwidths,
// and, for the channels to line up, equal channel counts and averaging
bool sideband_pair(const fxg_job_desc* d, int f, int g) {
  const fxg_freq& a = d->freqs[f];
  const fxg_freq& b = d->freqs[g];
  if (f == g || (a.lsb != 0) == (b.lsb != 0)) return false;
  const double la = a.freq_mhz - (a.lsb ? a.bw_mhz : 0.0), lb = b.freq_mhz - (b.lsb ? b.bw_mhz : 0.0);
  return la == lb && a.bw_mhz == b.bw_mhz && a.nchan == b.nchan && a.chans_to_avg == b.chans_to_avg;
}

// ---- phase cal ------------------------------------------------------------
// gcd of two values rounded to integers (mathutil.cpp:8-30)
long long gcd_ll(long long a, long long b) {
  if (a == 0 || b == 0) return std::max(a, b);
  for (;;) {
    a %= b;
    if (a == 0) return b;
    b %= a;
    if (b == 0) return a;
  }
}
long long gcd_d(double a, double b) { return gcd_ll((long long)std::floor(a + 0.5), (long long)std::floor(b + 0.5)); }

// tones from `offset` every `step` up to and including `bw`, a tone on both
// band edges counted once (PCal::calcNumTones, pcal.cpp:212-229)
int count_tones(double bw, double offset, double step) {
  int n = 0;
  if (bw <= 0.0 || step <= 0.0) return 0;
  double f = offset;
  while (f <= bw) { n++; f += step; }
  if (offset == 0.0 && f - step == bw) n--;
  return n;
}

// tones of one recorded frequency and the first tone's distance from the band
// edge in Hz, -1 when no tone falls inside (configuration.cpp:1771-1845)
void config_tones(const fxg_freq& fq, double interval, double base, int* ntones, int* offhz) {
  const double lo = fq.freq_mhz, bw = fq.bw_mhz;
  int n = 0;
  double tone = double(int(lo / interval)) * interval;
  double first;
  if (fq.lsb) {
    if (tone == lo) tone -= interval;
    if (tone >= lo - bw)
      while (tone - n * interval >= lo - bw) n++;
    first = 1e6 * tone - 1e6 * base;
    if (n > 0) *offhz = (int)(long)(1e6 * lo - first + 0.5);
  } else {
    if (tone <= lo) tone += interval;
    if (tone <= lo + bw)
      while (tone + n * interval <= lo + bw) n++;
    first = 1e6 * tone + 1e6 * base;
    if (n > 0) *offhz = (int)(long)(first - 1e6 * lo + 0.5);
  }
  if (n == 0) { n = 1; *offhz = -1; }
  *ntones = n;
}

// The extractor PCal::getNew picks with the 1 MHz minimum resolution Mode
// sets (mode.cpp:412-428; pcal.cpp:139-203), its bins and the DFT bin of each
// tone (-1: output zero).  kind 0: no tone (zeros); real samples 1: first
// tone at DC (pcal.cpp:333-350, 448-470), 2: implicit shift (:927-960,
// 1064-1104), 3: shifting (:498-540, 652-690); complex samples (sample rate
// = bandwidth) 4: implicit shift (:1107-1128, 1239-1285), 5: shifting
// (:707-760, 890-925).  Double sideband reads the DFT reversed (bin N-1-b).
struct PcalShape {
  int kind = 0;
  long long nbins = 0, period = 0;
  int ntones = 0;
  double dphi = 0.0;
  std::vector<int> tonebin;
};

PcalShape pcal_shape(double bw_hz, double spacing_hz, int offset_hz, bool cplx, bool dsb) {
  PcalShape p;
  const double fs = cplx ? bw_hz : 2.0 * bw_hz, minres = 1e6;
  if (offset_hz == -1) return p;
  long long No = (long long)(fs / gcd_d(fs, (double)offset_hz));
  const long long Np = (long long)(fs / gcd_d(fs, spacing_hz));
  const int Nt = count_tones(bw_hz, (double)offset_hz, spacing_hz);
  while (fs / (2 * No) > minres) No *= 2;
  if (Nt == 0) return p;
  if (cplx) {
    const bool implicit = No % Np == 0;
    p.kind = implicit ? 4 : 5;
    p.nbins = (long long)(fs / gcd_d(fs, implicit ? (double)offset_hz : spacing_hz));
    while (fs / (2 * p.nbins) > minres) p.nbins *= 2;
    p.ntones = Nt;
    if (!implicit) {
      // the mixer's values repeat with the offset's period (the reference's
      // table runs over the lcm of that and N, a whole number of periods)
      p.period = (long long)(fs / gcd_d(fs, offset_hz));
      p.dphi = kTwoPi * (-offset_hz / fs);
    }
    const long long step = (long long)std::floor(double(p.nbins) * (spacing_hz / fs));
    const long long off = implicit ? (long long)std::floor(double(p.nbins) * (offset_hz / fs)) : 0;
    for (int i = 0; i < p.ntones; i++) {
      const long long idx = off + i * step;
      // the reference stops copying at the first bin past the DFT; zero
      p.tonebin.push_back(idx >= p.nbins ? -1 : (int)(dsb ? p.nbins - 1 - idx : idx));
    }
    return p;
  }
  if (offset_hz == 0) {
    p.kind = 1;
    p.nbins = (long long)(fs / gcd_d(fs, spacing_hz));
    p.ntones = count_tones(bw_hz, 0.0, spacing_hz) - 1;
    const long long step = (long long)(spacing_hz / gcd_d(fs, spacing_hz));
    for (int i = 0; i < p.ntones; i++) p.tonebin.push_back((int)((i + 1) * step));
  } else if (No % Np == 0) {
    p.kind = 2;
    p.nbins = (long long)(fs / gcd_d(fs, offset_hz));
    while (fs / (2 * p.nbins) > minres) p.nbins *= 2;
    p.ntones = Nt;
    const long long step = (long long)std::floor(double(p.nbins) * (spacing_hz / fs));
    const long long off = (long long)std::floor(double(p.nbins) * (offset_hz / fs));
    for (int i = 0; i < p.ntones; i++) {
      long long idx = off + i * step;
      p.tonebin.push_back(idx >= p.nbins ? -1 : (int)idx);
    }
  } else {
    p.kind = 3;
    p.nbins = (long long)(fs / gcd_d(fs, spacing_hz));
    p.ntones = Nt;
    p.period = (long long)(fs / gcd_d(fs, offset_hz));
    p.dphi = kTwoPi * (-offset_hz / fs);
    const long long step = (long long)std::floor(p.nbins * (spacing_hz / fs));
    for (int i = 0; i < p.ntones; i++) {
      long long idx = i * step;
      // the reference stops copying here and leaves the rest untouched; zero
      p.tonebin.push_back(idx >= p.nbins ? -1 : (int)idx);
    }
  }
  return p;
}

#define REQUIRE(cond, code, msg) do { if (!(cond)) { *err = (msg); return (code); } } while (0)

int check_desc(const fxg_job_desc* d, std::string* err) {
  REQUIRE(d && d->freqs && d->datastreams, FXG_EINVAL, "null descriptor table");
  REQUIRE(d->nbaselines == 0 || d->baselines, FXG_EINVAL, "null baseline table");
  REQUIRE(d->nfreqs > 0 && d->nfreqs <= FXG_MAX_FREQS, FXG_EINVAL, "nfreqs out of range");
  REQUIRE(d->ndatastreams > 0, FXG_EINVAL, "ndatastreams must be positive");
  REQUIRE(d->nbaselines >= 0, FXG_EINVAL, "nbaselines negative");
  REQUIRE(d->blockspersend > 0, FXG_EINVAL, "blockspersend must be positive");
  REQUIRE(d->fringe_order >= 0 && d->fringe_order <= 2, FXG_EUNSUPPORTED, "fringe rotation order must be 0, 1 or 2");
  REQUIRE(d->xmac_precision >= FXG_XMAC_FP32 && d->xmac_precision <= FXG_XMAC_FP8, FXG_EINVAL,
          "xmac_precision");
  for (int f = 0; f < d->nfreqs; f++) {
    REQUIRE(d->freqs[f].nchan > 0, FXG_EINVAL, "freq nchan must be positive");
    REQUIRE(d->freqs[f].bw_mhz > 0, FXG_EINVAL, "freq bandwidth must be positive");
    REQUIRE(d->freqs[f].chans_to_avg >= 1 && d->freqs[f].nchan % d->freqs[f].chans_to_avg == 0, FXG_EINVAL,
            "chans_to_avg must divide nchan");
  }
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    REQUIRE(ds.format == FXG_FORMAT_VDIF || ds.format == FXG_FORMAT_MARK5B || ds.format == FXG_FORMAT_INTERLACEDVDIF,
            FXG_EINVAL, "format");
    // INTERLACEDVDIF: the VDIF decoders plus Mk5Mode's EDV4 per-band blanking
    // of real samples (blank_vdif_EDV4, format_vdif.c:4699-4775)
    REQUIRE(!(ds.format == FXG_FORMAT_INTERLACEDVDIF && ds.sampling != FXG_SAMPLING_REAL), FXG_EUNSUPPORTED,
            "INTERLACEDVDIF with complex sampling");
    REQUIRE(!(ds.format == FXG_FORMAT_INTERLACEDVDIF && ds.headerbytes != 32), FXG_EINVAL,
            "INTERLACEDVDIF frames carry the 32-byte EDV4 header");
    REQUIRE(ds.nrecfreqs > 0 && ds.nrecfreqs <= FXG_MAX_RECFREQS, FXG_EINVAL, "nrecfreqs");
    REQUIRE(ds.nrecbands > 0 && ds.nrecbands <= FXG_MAX_BANDS, FXG_EINVAL, "nrecbands");
    REQUIRE(power_of_two(ds.nrecbands), FXG_EUNSUPPORTED, "nrecbands must be a power of two");
    REQUIRE(ds.nzoombands >= 0 && ds.nzoombands <= FXG_MAX_ZOOMBANDS, FXG_EINVAL, "nzoombands");
    REQUIRE(ds.sampling == FXG_SAMPLING_REAL || ds.sampling == FXG_SAMPLING_COMPLEX, FXG_EINVAL, "sampling");
    REQUIRE(ds.sampling == d->datastreams[0].sampling, FXG_EUNSUPPORTED,
            "all datastreams of a job must share real or complex sampling (one FFT length per job)");
    if (ds.sampling == FXG_SAMPLING_COMPLEX) {
      REQUIRE(ds.format == FXG_FORMAT_VDIF, FXG_EUNSUPPORTED, "complex sampling is VDIF only");
      REQUIRE(d->fringe_order != 0, FXG_EUNSUPPORTED,
              "post-F fringe rotation is not supported for complex data (mode.cpp:1155-1158)");
    }
    if (ds.format != FXG_FORMAT_MARK5B) {
      REQUIRE(ds.nbits == 2 || ds.nbits == 8, FXG_EUNSUPPORTED, "VDIF nbits must be 2 or 8");
    } else {
      REQUIRE(ds.nbits == 2, FXG_EUNSUPPORTED, "Mark5B nbits must be 2");
      REQUIRE(ds.headerbytes == 16 && ds.framebytes == 10016, FXG_EINVAL, "Mark5B frame is 16+10000 bytes");
      REQUIRE(ds.nrecbands * ds.nbits <= 32, FXG_EUNSUPPORTED, "Mark5B bitstreams <= 32");
    }
    REQUIRE(ds.headerbytes >= 0 && ds.framebytes > ds.headerbytes, FXG_EINVAL, "frame sizes");
    REQUIRE((ds.framebytes - ds.headerbytes) % 8 == 0, FXG_EINVAL, "payload must be a multiple of 8 bytes");
    for (int i = 0; i < ds.nrecfreqs; i++) {
      REQUIRE(ds.recfreq_index[i] >= 0 && ds.recfreq_index[i] < d->nfreqs, FXG_EINVAL, "recfreq_index");
    }
    for (int b = 0; b < ds.nrecbands; b++)
      REQUIRE(ds.recband_freq[b] >= 0 && ds.recband_freq[b] < ds.nrecfreqs, FXG_EINVAL, "recband_freq");
    for (int z = 0; z < ds.nzoombands; z++) {
      REQUIRE(ds.zoomband_freqindex[z] >= 0 && ds.zoomband_freqindex[z] < d->nfreqs, FXG_EINVAL, "zoom freq");
      REQUIRE(ds.zoomband_parent[z] >= 0 && ds.zoomband_parent[z] < ds.nrecbands, FXG_EINVAL, "zoom parent");
      int pc = d->freqs[ds.recfreq_index[ds.recband_freq[ds.zoomband_parent[z]]]].nchan;
      REQUIRE(ds.zoomband_choffset[z] >= 0 &&
              ds.zoomband_choffset[z] + d->freqs[ds.zoomband_freqindex[z]].nchan <= pc,
              FXG_EINVAL, "zoom band outside its parent");
    }
  }
  for (int b = 0; b < d->nbaselines; b++) {
    const fxg_baseline& bl = d->baselines[b];
    REQUIRE(bl.ds1 >= 0 && bl.ds1 < d->ndatastreams && bl.ds2 >= 0 && bl.ds2 < d->ndatastreams,
            FXG_EINVAL, "baseline datastream index");
    REQUIRE(bl.nfreqs >= 0 && bl.nfreqs <= FXG_MAX_BLFREQS, FXG_EINVAL, "baseline nfreqs");
    const fxg_datastream& a = d->datastreams[bl.ds1];
    const fxg_datastream& c = d->datastreams[bl.ds2];
    for (int f = 0; f < bl.nfreqs; f++) {
      int fi = bl.freq_index[f];
      REQUIRE(fi >= 0 && fi < d->nfreqs, FXG_EINVAL, "baseline freq index");
      REQUIRE(bl.npolproducts[f] > 0 && bl.npolproducts[f] <= FXG_MAX_POLPRODUCTS, FXG_EINVAL, "npolproducts");
      for (int p = 0; p < bl.npolproducts[f]; p++) {
        int b1 = bl.band1[f][p], b2 = bl.band2[f][p];
        REQUIRE(b1 >= 0 && b1 < a.nrecbands + a.nzoombands && b2 >= 0 && b2 < c.nrecbands + c.nzoombands,
                FXG_EINVAL, "baseline band index");
        const int f1 = freq_of_band(a, b1), f2 = freq_of_band(c, b2);
        REQUIRE((f1 == fi || sideband_pair(d, f1, fi)) && (f2 == fi || sideband_pair(d, f2, fi)), FXG_EINVAL,
                "baseline band carries neither the baseline frequency nor its opposite-sideband equivalent");
      }
    }
  }
  return 0;
}

}  // namespace

int build_plan(const fxg_job_desc* d, Plan* p, std::string* err) {
  int rc = check_desc(d, err);
  if (rc) return rc;
  Plan& P = *p;
  P = Plan();
  P.nfreqs = d->nfreqs; P.nds = d->ndatastreams; P.nbl = d->nbaselines;

  // ---- used / equivalent-used frequencies, per-baseline local freq index
  P.freqused.assign(d->nfreqs, 0);
  P.equivused.assign(d->nfreqs, 0);
  P.bllocal.assign(d->nbaselines, std::vector<int>(d->nfreqs, -1));
  int maxprod = 0;
  for (int b = 0; b < d->nbaselines; b++)
    for (int f = 0; f < d->baselines[b].nfreqs; f++) {
      int fi = d->baselines[b].freq_index[f];
      P.freqused[fi] = 1;
      P.bllocal[b][fi] = f;
      maxprod = std::max(maxprod, d->baselines[b].npolproducts[f]);
    }
  for (int f = 0; f < d->nfreqs; f++) {
    if (P.freqused[f]) continue;
    const fxg_freq& x = d->freqs[f];
    double xlo = x.freq_mhz - (x.lsb ? x.bw_mhz : 0.0);
    for (int g = 0; g < d->nfreqs && !P.equivused[f]; g++) {
      const fxg_freq& y = d->freqs[g];
      double ylo = y.freq_mhz - (y.lsb ? y.bw_mhz : 0.0);
      // signed comparisons, as the reference makes them
      if (P.freqused[g] && x.bw_mhz - y.bw_mhz < kTiny && xlo - ylo < kTiny &&
          x.nchan == y.nchan && x.chans_to_avg == y.chans_to_avg)
        P.equivused[f] = 1;
    }
  }
  P.bandsperac = maxprod > 2 ? 2 : 1;

  // ---- strides
  P.arraystride.resize(d->ndatastreams);
  for (int s = 0; s < d->ndatastreams; s++) {
    int nchan = d->freqs[d->datastreams[s].recfreq_index[0]].nchan;
    P.arraystride[s] = d->arraystridelen > 0 ? d->arraystridelen : stride_for((unsigned)nchan);
  }
  P.xmacstride = d->xmacstridelen > 0 ? d->xmacstridelen : xmac_stride_for(d);
  P.rotatestride = stride_for((unsigned)P.xmacstride);
  for (int f = 0; f < d->nfreqs; f++)
    REQUIRE(!P.freqused[f] || d->freqs[f].nchan % P.xmacstride == 0, FXG_EINVAL,
            "xmac stride must divide every correlated channel count");

  // ---- result layout
  long long at = 0;
  P.bloff.assign(d->nfreqs, std::vector<long long>(d->nbaselines, -1));
  P.bwoff.assign(d->nfreqs, std::vector<long long>(d->nbaselines, -1));
  for (int f = 0; f < d->nfreqs; f++) {
    if (!P.freqused[f]) continue;
    for (int b = 0; b < d->nbaselines; b++) {
      int lf = P.bllocal[b][f];
      if (lf < 0) continue;
      P.bloff[f][b] = at;
      at += (long long)d->baselines[b].npolproducts[lf] * d->freqs[f].nchan / d->freqs[f].chans_to_avg;
    }
  }
  for (int f = 0; f < d->nfreqs; f++) {
    if (!P.freqused[f]) continue;
    for (int b = 0; b < d->nbaselines; b++) {
      int lf = P.bllocal[b][f];
      if (lf < 0) continue;
      P.bwoff[f][b] = at;
      at += (d->baselines[b].npolproducts[lf] + 1) / 2;
    }
  }
  auto ac_wanted = [&](int fi) { return P.freqused[fi] || P.equivused[fi]; };
  P.acoff.resize(d->ndatastreams);
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    P.acoff[s] = at;
    for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++)
      if (ac_wanted(freq_of_band(ds, k)))
        at += (long long)P.bandsperac * d->freqs[freq_of_band(ds, k)].nchan / d->freqs[freq_of_band(ds, k)].chans_to_avg;
  }
  P.acwoff.resize(d->ndatastreams);
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    P.acwoff[s] = at;
    int n = 0;
    for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++)
      if (ac_wanted(freq_of_band(ds, k))) n += P.bandsperac;
    at += std::max(1, (n + 1) / 2);
  }
  // phase-cal tones per (datastream, recorded freq) and their result region
  P.pcaloff.assign(d->ndatastreams, 0);
  P.pcal_ntones.assign(d->ndatastreams, {});
  P.pcal_offhz.assign(d->ndatastreams, {});
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    P.pcaloff[s] = at;
    if (!(ds.pcal_interval_mhz > 0)) continue;
    P.pcal_ntones[s].resize(ds.nrecfreqs);
    P.pcal_offhz[s].resize(ds.nrecfreqs);
    for (int i = 0; i < ds.nrecfreqs; i++)
      config_tones(d->freqs[ds.recfreq_index[i]], ds.pcal_interval_mhz, ds.pcal_base_mhz,
                   &P.pcal_ntones[s][i], &P.pcal_offhz[s][i]);
    for (int k = 0; k < ds.nrecbands; k++) at += P.pcal_ntones[s][ds.recband_freq[k]];
  }

  // ---- per-datastream device constants; one FFT length per job
  P.C = d->freqs[d->datastreams[0].recfreq_index[0]].nchan;
  P.L = d->datastreams[0].sampling == FXG_SAMPLING_COMPLEX ? P.C : 2 * P.C;
  int band = 0;
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    const fxg_freq& f0 = d->freqs[ds.recfreq_index[0]];
    for (int i = 0; i < ds.nrecfreqs; i++) {
      const fxg_freq& fi = d->freqs[ds.recfreq_index[i]];
      REQUIRE(fi.nchan == f0.nchan && fi.bw_mhz == f0.bw_mhz, FXG_EUNSUPPORTED,
              "recorded freqs of a datastream must share nchan and bandwidth");
    }
    if (f0.nchan != P.C) P.uniform = false;
    DevStation st{};
    st.edv4 = ds.format == FXG_FORMAT_INTERLACEDVDIF ? 1 : 0;
    if (st.edv4) P.any_edv4 = true;
    st.format = st.edv4 ? FXG_FORMAT_VDIF : ds.format; st.nbits = ds.nbits; st.nchan = ds.nrecbands;
    st.framebytes = ds.framebytes; st.headerbytes = ds.headerbytes;
    st.databytes = ds.framebytes - ds.headerbytes; st.blanker = ds.blanker;
    st.cplx = ds.sampling == FXG_SAMPLING_COMPLEX ? 1 : 0;
    st.unitbits = st.cplx ? 2 * ds.nbits : ds.nbits;
    int bitsper = ds.nrecbands * st.unitbits;
    st.framesamples = (int)((long long)st.databytes * 8 / bitsper);
    st.granularity = bitsper >= 8 ? 1 : 8 / bitsper;
    st.groupbytes = bitsper >= 8 ? bitsper / 8 : 1;
    st.groupsamples = bitsper >= 8 ? 1 : 8 / bitsper;
    // decoders that stop at a blank group: real 8-bit with 2..4 bands
    // (format_vdif.c:1837-2020), complex 8-bit with any (:2829-3054)
    st.quirk8 = st.format == FXG_FORMAT_VDIF && ds.nbits == 8 &&
                (st.cplx || (ds.nrecbands >= 2 && ds.nrecbands <= 4));
    // complex: fftchannels = nchan, sample time 1/bw (mode.cpp:82-87)
    st.C = f0.nchan; st.L = st.cplx ? f0.nchan : 2 * f0.nchan; st.s = P.arraystride[s];
    REQUIRE(st.C % st.s == 0, FXG_EINVAL, "array stride must divide nchan");
    st.nfrsteps = st.L / st.s; st.nfracsteps = st.C / st.s;
    st.slog = -1;
    if (power_of_two(st.s)) { st.slog = 0; while ((1 << st.slog) < st.s) st.slog++; }
    REQUIRE(st.databytes % ds.nrecbands == 0, FXG_EUNSUPPORTED, "payload must split evenly over the bands");
    st.bpf = st.databytes / ds.nrecbands;
    st.order = d->fringe_order;
    st.bpbnum = bitsper / 8;
    st.bpbden = 1;
    if (st.bpbnum == 0) { st.bpbnum = 1; st.bpbden = 8 / bitsper; }
    st.bw = f0.bw_mhz;
    st.sampletime = 1.0 / (2.0 * st.bw);
    if (st.cplx) st.sampletime *= 2.0;
    const bool dsb = st.cplx && ds.complex_dsb != 0;
    st.fractionalLo = 0; st.deltapol = 0;
    for (int i = 0; i < ds.nrecfreqs; i++) {
      const fxg_freq& fi = d->freqs[ds.recfreq_index[i]];
      double lo = fi.freq_mhz;
      if (dsb) lo += fi.lsb ? -fi.bw_mhz / 2.0 : fi.bw_mhz / 2.0;      // mode.cpp:95-105
      if (std::fabs(lo - (int)(lo + 0.5)) > kTiny) st.fractionalLo = 1;
      if (ds.recfreq_clockoffsetdelta_us[i] != 0.0) st.deltapol = 1;
    }
    st.band0 = band;
    if (st.format == FXG_FORMAT_VDIF) { st.lut[0] = -kHiMag; st.lut[1] = -1.0f; st.lut[2] = 1.0f; st.lut[3] = kHiMag; }
    else { st.lut[0] = -kHiMag; st.lut[1] = 1.0f; st.lut[2] = -1.0f; st.lut[3] = kHiMag; }
    P.st.push_back(st);
    for (int k = 0; k < ds.nrecbands; k++) {
      DevBand b{};
      int lf = ds.recband_freq[k];
      const fxg_freq& fq = d->freqs[ds.recfreq_index[lf]];
      b.station = s; b.band = k; b.localfreq = lf; b.lsb = fq.lsb ? 1 : 0;
      b.useB = (st.deltapol && ds.recband_pol[k] != 'R') ? 1 : 0;
      b.dsb = dsb ? 1 : 0;
      // the fringe rotator's LO (mode.cpp:886-908): the band edge; the band
      // centre for double-sideband complex data; negative for complex LSB
      b.lofreq = fq.freq_mhz;
      if (dsb) b.lofreq += fq.lsb ? -fq.bw_mhz / 2.0 : fq.bw_mhz / 2.0;
      if (st.cplx && fq.lsb) b.lofreq = -b.lofreq;
      b.clock = ds.recfreq_clockoffset_us[lf];
      b.clockdelta = ds.recfreq_clockoffsetdelta_us[lf];
      b.looffset = ds.recfreq_looffset_hz[lf];
      // polarisation phase offset (mode.cpp:435-460, 1341-1353): with two or
      // more bands of the frequency, the second of the first two if the first
      // is R, else the first, is turned by phasecorrA = e^{i offset/2}
      b.polre = 1.0f; b.polim = 0.0f;
      {
        int i0 = -1, i1 = -1;
        for (int q = 0; q < ds.nrecbands; q++)
          if (ds.recband_freq[q] == lf) { if (i0 < 0) i0 = q; else if (i1 < 0) i1 = q; }
        bool anyoff = false;
        for (int i = 0; i < ds.nrecfreqs; i++) anyoff |= ds.recfreq_phaseoffset_deg[i] != 0.0;
        const int lcp = i1 < 0 ? -1 : (ds.recband_pol[i0] == 'R' ? i1 : i0);
        if (anyoff && k == lcp) {
          const double degphase = -ds.recfreq_phaseoffset_deg[lf] / 2;
          b.polre = (float)std::cos(degphase * M_PI / 180.0);
          b.polim = (float)-std::sin(degphase * M_PI / 180.0);
        }
      }
      P.bands.push_back(b);
      band++;
    }
  }
  P.nbands_total = band;

  // ---- phase-cal device tables
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    if (P.pcal_ntones[s].empty()) continue;
    const double bwhz = 1e6 * d->freqs[ds.recfreq_index[0]].bw_mhz;
    const bool csamp = ds.sampling == FXG_SAMPLING_COMPLEX;
    const double fsr = csamp ? bwhz : 2 * bwhz;
    long long at3 = P.pcaloff[s];
    for (int k = 0; k < ds.nrecbands; k++) {
      const int lf = ds.recband_freq[k];
      const int nt = P.pcal_ntones[s][lf];
      REQUIRE(P.pcal_offhz[s][lf] >= -1, FXG_EUNSUPPORTED,
              "phase cal: the tone base puts the first tone outside the band");
      PcalShape sh = pcal_shape(bwhz, 1e6 * ds.pcal_interval_mhz, P.pcal_offhz[s][lf], csamp,
                                csamp && ds.complex_dsb != 0);
      const long long resoff = at3;
      at3 += nt;
      if (sh.kind == 0) continue;                       // no tone in the band: zeros
      REQUIRE(sh.ntones == nt, FXG_EUNSUPPORTED,
              "phase cal: the extractor's tone count differs from the configuration's");
      REQUIRE(std::fabs(fsr - std::floor(fsr + 0.5)) < 1e-6, FXG_EUNSUPPORTED,
              "phase cal needs an integral sample rate in Hz");
      REQUIRE(sh.nbins * (sh.kind >= 3 ? 2 : 1) <= kPcalMaxFloats, FXG_EUNSUPPORTED,
              "phase cal: too many accumulation bins for one workgroup");
      REQUIRE(sh.period <= (1 << 22), FXG_EUNSUPPORTED, "phase cal: rotator period too long");
      DevPcal pc{};
      pc.gband = P.st[s].band0 + k;
      pc.nbins = (int)sh.nbins;
      pc.cplx = sh.kind >= 3;
      pc.rot = sh.kind == 3 || sh.kind == 5;
      pc.csamp = csamp;
      pc.period = (int)sh.period;
      pc.ntones = nt;
      pc.toneoff = (int)P.pcal_tonebin.size();
      P.pcal_tonebin.insert(P.pcal_tonebin.end(), sh.tonebin.begin(), sh.tonebin.end());
      pc.resoff = resoff;
      if (pc.rot) {
        // the shifting extractors' mixer, (f32) of the FP64 value (pcal.cpp:530-535, 751-756)
        pc.rotoff = (long long)P.pcal_aux.size() / 2;
        for (long long m = 0; m < sh.period; m++) {
          const double arg = sh.dphi * double(m);
          P.pcal_aux.push_back((float)std::cos(arg));
          P.pcal_aux.push_back((float)std::sin(arg));
        }
      }
      pc.twoff = (long long)P.pcal_aux.size() / 2;
      for (long long m = 0; m < sh.nbins; m++) {
        const double arg = -kTwoPi * double(m) / double(sh.nbins);
        P.pcal_aux.push_back((float)std::cos(arg));
        P.pcal_aux.push_back((float)std::sin(arg));
      }
      P.pcal_maxtones = std::max(P.pcal_maxtones, nt);
      P.pcal.push_back(pc);
    }
  }
  if (!P.pcal.empty()) {
    // enough (band, block group) workgroups to cover the chip
    const int nb = (int)P.pcal.size();
    P.pcal_groups = std::max(1, std::min(d->blockspersend, (1024 + nb - 1) / nb));
    long long part = 0;
    for (DevPcal& pc : P.pcal) {
      pc.partoff = part;
      part += (long long)P.pcal_groups * pc.nbins * (pc.cplx ? 2 : 1);
    }
    P.pcal_partfloats = part;
  }

  // ---- XMAC frequencies: every frequency whose products or autocorrelations
  //      are output (zoom autocorrelations read their parent's accumulation)
  std::vector<char> need(d->nfreqs, 0);
  for (int f = 0; f < d->nfreqs; f++) need[f] = P.freqused[f];
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
      int fk = freq_of_band(ds, k);
      if (!ac_wanted(fk)) continue;
      need[ds.recfreq_index[ds.recband_freq[parent_of_band(ds, k)]]] = 1;
    }
  }
  P.xf_of_freq.assign(d->nfreqs, -1);
  long long voff = 0;
  for (int f = 0; f < d->nfreqs; f++) {
    if (!need[f]) continue;
    XmacFreq x;
    x.freq = f; x.nchan = d->freqs[f].nchan;
    // bands of the opposite sideband that baselines of f correlate against it
    std::vector<std::vector<char>> cross(d->ndatastreams);
    for (int s = 0; s < d->ndatastreams; s++)
      cross[s].assign(d->datastreams[s].nrecbands + d->datastreams[s].nzoombands, 0);
    for (int b = 0; b < d->nbaselines; b++) {
      const int lf = P.bllocal[b][f];
      if (lf < 0) continue;
      const fxg_baseline& bl = d->baselines[b];
      for (int pp = 0; pp < bl.npolproducts[lf]; pp++) {
        cross[bl.ds1][bl.band1[lf][pp]] = 1;
        cross[bl.ds2][bl.band2[lf][pp]] = 1;
      }
    }
    for (int s = 0; s < d->ndatastreams; s++) {
      const fxg_datastream& ds = d->datastreams[s];
      for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
        if (freq_of_band(ds, k) != f && !cross[s][k]) continue;
        x.in_station.push_back(s);
        x.in_band.push_back(parent_of_band(ds, k));
        x.in_total.push_back(k);
        x.in_off.push_back(choff_of_band(ds, k));   // completed with band offsets at allocation
      }
    }
    if (x.in_station.empty()) continue;
    x.voff = voff;
    voff += (long long)x.in_station.size() * x.in_station.size() * x.nchan;
    P.xf_of_freq[f] = (int)P.xf.size();
    P.xf.push_back(x);
  }
  P.vsize = voff;

  // ---- scatter segments
  auto input_index = [&](int xfi, int s, int totalband) {
    const XmacFreq& x = P.xf[xfi];
    for (size_t i = 0; i < x.in_station.size(); i++)
      if (x.in_station[i] == s && x.in_total[i] == totalband) return (int)i;
    return -1;
  };
  bool seg_ok = true;
  // nchan output channels, each the mean of avg consecutive V channels from choff
  auto vis_seg = [&](long long resoff, int xfi, int a, int b, int nchan, int choff, int avg) {
    if (xfi < 0 || a < 0 || b < 0) { seg_ok = false; return; }
    const XmacFreq& x = P.xf[xfi];
    int nin = (int)x.in_station.size();
    DevSeg g{};
    g.kind = SEG_VIS; g.resoff = resoff; g.nchan = nchan; g.avg = avg;
    // V is accumulated for a <= b tiles and in full inside diagonal tiles;
    // read the upper triangle and conjugate when the order is swapped
    if (a <= b) { g.voff = x.voff + ((long long)a * nin + b) * x.nchan + choff; g.conj = 0; }
    else        { g.voff = x.voff + ((long long)b * nin + a) * x.nchan + choff; g.conj = 1; }
    P.segs.push_back(g);
    P.maxseglen = std::max(P.maxseglen, nchan);
  };
  for (int f = 0; f < d->nfreqs; f++) {
    if (!P.freqused[f]) continue;
    int xfi = P.xf_of_freq[f];
    for (int b = 0; b < d->nbaselines; b++) {
      int lf = P.bllocal[b][f];
      if (lf < 0) continue;
      const fxg_baseline& bl = d->baselines[b];
      for (int pp = 0; pp < bl.npolproducts[lf]; pp++) {
        int ia = input_index(xfi, bl.ds1, bl.band1[lf][pp]);
        int ib = input_index(xfi, bl.ds2, bl.band2[lf][pp]);
        // channel averaging of the products (core.cpp:1788-1802)
        const int avg = d->freqs[f].chans_to_avg, nout = d->freqs[f].nchan / avg;
        vis_seg(P.bloff[f][b] + (long long)pp * nout, xfi, ia, ib, nout, 0, avg);
      }
    }
  }
  // weight table entries: one per (datastream, band) pairing that a weight
  // of the result reads; the band is -1 (the block's data weight) except on
  // INTERLACEDVDIF datastreams, whose recorded bands carry their own weights
  // (Mode::getDataWeight, mode.h:168), so that a job without them has one
  // entry per baseline and one per datastream
  P.wterms.clear();
  std::map<std::array<int, 4>, long long> wterm_index;
  auto wterm = [&](int d1, int b1, int d2, int b2) -> long long {
    if (!P.st[d1].edv4) b1 = -1;
    if (d2 >= 0 && !P.st[d2].edv4) b2 = -1;
    if (d2 < 0) b2 = -1;
    const std::array<int, 4> key{d1, b1, d2, b2};
    auto it = wterm_index.find(key);
    if (it != wterm_index.end()) return it->second;
    P.wterms.push_back(DevWTerm{d1, b1, d2, b2});
    wterm_index[key] = (long long)P.wterms.size() - 1;
    return (long long)P.wterms.size() - 1;
  };
  auto record_band = [&](int s, int band) {
    const fxg_datastream& ds = d->datastreams[s];
    return band < ds.nrecbands ? band : ds.zoomband_parent[band - ds.nrecbands];
  };
  for (int b = 0; b < d->nbaselines; b++) wterm(d->baselines[b].ds1, -1, d->baselines[b].ds2, -1);
  for (int s = 0; s < d->ndatastreams; s++) wterm(s, -1, -1, -1);
  // baseline weights: floats, the product of the two record bands' weights
  // summed over the blocks (core.cpp:1028-1037)
  for (int f = 0; f < d->nfreqs; f++) {
    if (!P.freqused[f]) continue;
    for (int b = 0; b < d->nbaselines; b++) {
      int lf = P.bllocal[b][f];
      if (lf < 0) continue;
      const fxg_baseline& bl = d->baselines[b];
      for (int pp = 0; pp < bl.npolproducts[lf]; pp++) {
        DevSeg g{}; g.kind = SEG_FLOATS; g.resoff = 2 * P.bwoff[f][b] + pp;
        g.voff = wterm(bl.ds1, record_band(bl.ds1, bl.band1[lf][pp]), bl.ds2, record_band(bl.ds2, bl.band2[lf][pp]));
        g.nchan = 1;
        P.segs.push_back(g);
      }
    }
  }
  // autocorrelations (Mode::process accumulation, mode.cpp:1292-1375, copied
  // out by averageAndSendAutocorrs, core.cpp:1299-1330)
  for (int s = 0; s < d->ndatastreams; s++) {
    const fxg_datastream& ds = d->datastreams[s];
    long long at2 = P.acoff[s];
    long long wat = 2 * P.acwoff[s];
    for (int pass = 0; pass < P.bandsperac; pass++) {
      for (int k = 0; k < ds.nrecbands + ds.nzoombands; k++) {
        int fk = freq_of_band(ds, k);
        if (!ac_wanted(fk)) continue;
        // Mode::averageFrequency averages with the datastream's first recorded
        // frequency's factor (configuration.cpp:962, mode.cpp:1399-1424); a zoom
        // band then starts at choffset/avg of its averaged parent (mode.cpp:396)
        const int savg = d->freqs[ds.recfreq_index[0]].chans_to_avg;
        int nch = d->freqs[fk].nchan / d->freqs[fk].chans_to_avg;
        int parent = parent_of_band(ds, k);
        int pf = ds.recfreq_index[ds.recband_freq[parent]];
        int xfi = P.xf_of_freq[pf];
        int choff = (choff_of_band(ds, k) / savg) * savg;
        if (pass == 0) {
          int a = input_index(xfi, s, parent);
          vis_seg(at2, xfi, a, a, nch, choff, savg);
        } else if (d->crosspol_autocorrs) {
          // the first two recorded bands of the parent's local frequency
          int lf = ds.recband_freq[parent];
          int i0 = -1, i1 = -1;
          for (int q = 0; q < ds.nrecbands; q++)
            if (ds.recband_freq[q] == lf) { if (i0 < 0) i0 = q; else if (i1 < 0) i1 = q; }
          if (i1 >= 0 && (parent == i0 || parent == i1)) {
            int other = parent == i0 ? i1 : i0;
            vis_seg(at2, xfi, input_index(xfi, s, parent), input_index(xfi, s, other), nch, choff, savg);
          }
        }
        at2 += nch;
        // AC weights: getWeight(false, parent) for both passes (core.cpp:1331-1366)
        DevSeg w{}; w.kind = SEG_FLOATS; w.resoff = wat++; w.voff = wterm(s, parent, -1, -1); w.nchan = 1;
        if (pass == 0 || d->crosspol_autocorrs) P.segs.push_back(w);
      }
    }
  }

  REQUIRE(seg_ok, FXG_EINVAL, "a result product names a band that no XMAC frequency holds");
  if (d->xmac_precision == FXG_XMAC_FP8)
    for (const XmacFreq& x : P.xf)
      REQUIRE(x.in_station.size() <= 64 && x.nchan % 16 == 0, FXG_EUNSUPPORTED,
              "fp8 cross-multiply needs at most 64 inputs and a multiple of 16 channels per frequency");
  P.info.result_length = at;
  P.info.arraystridelen = P.arraystride[0];
  P.info.xmacstridelen = P.xmacstride;
  P.info.rotatestridelen = P.rotatestride;
  P.info.bandsperautocorr = P.bandsperac;
  P.info.fftchannels = P.L;

  // ---- spectra chunking: keep the spectra of at most ~16 GiB resident
  long long per_block = (long long)P.nbands_total * P.C * spectrum_bytes(d->xmac_precision);
  long long budget = 16LL << 30;
  int chunk = d->max_chunk_blocks > 0 ? d->max_chunk_blocks : (int)std::max(1LL, budget / per_block);
  chunk = std::min(chunk, d->blockspersend);
  P.chunk_blocks = chunk;
  P.info.chunk_blocks = chunk;
  // FFTs above 16384 points: a scratch of about 160 MiB takes zcap blocks of
  // the chunk at a time, with their rotator tables; at that size the second
  // kernel reads back what the first wrote from the 256 MB Infinity Cache
  // rather than from HBM (measured: the 65536-point job of the uneven mix
  // 54.0 -> 51.5 ms per subint against a 2 GiB scratch)
  if (P.L > 16384) {
    P.zcap = (int)std::max(1LL, std::min<long long>(chunk, (160LL << 20) / ((long long)P.nbands_total * P.L * 8)));
    for (const DevStation& st : P.st) P.rtabn = std::max(P.rtabn, 3 * st.s + st.nfrsteps + st.nfracsteps);
  }
  // partial matrices for XMAC launches with few channel groups (block split):
  // up to 512 workgroups' worth, at most 256 MiB
  for (const XmacFreq& x : P.xf) {
    const long long nin = (long long)x.in_station.size();
    const long long vsz = nin * nin * x.nchan;
    const long long wgx = std::max(1, x.nchan / 16);   // the widest channel group used
    if (wgx >= 256 || chunk < 16) continue;
    const long long rows = std::min<long long>((512 + wgx - 1) / wgx, chunk / 8);
    P.vpart_elems = std::max(P.vpart_elems, std::min(rows * vsz, (256LL << 20) / 8 / std::max(1LL, vsz) * vsz));
  }
  P.info.device_bytes = plan_device_bytes(P, d);
  return 0;
}

int64_t plan_device_bytes(const Plan& p, const fxg_job_desc* d) {
  long long spec = (long long)p.nbands_total * p.C * p.chunk_blocks * spectrum_bytes(d->xmac_precision);
  if (d->xmac_precision == FXG_XMAC_FP8) spec += (long long)p.nbands_total * p.C;   // the scale exponents
  long long v = p.vsize * 8;
  long long blocks = (long long)p.nds * d->blockspersend * (long long)sizeof(DevBlock);
  long long res = p.info.result_length * 8;
  long long tables = (long long)p.segs.size() * sizeof(DevSeg) + p.bands.size() * sizeof(DevBand) + (1 << 20);
  long long z = p.L > 16384 ? (long long)p.nbands_total * p.zcap * ((long long)p.L + p.rtabn) * 8 : 0;
  long long vpart = p.vpart_elems * 8;
  long long pcal = p.pcal_partfloats * 4 + (long long)p.pcal_aux.size() * 4 +
                   (long long)p.pcal.size() * sizeof(DevPcal) + (long long)p.pcal_tonebin.size() * 4;
  // corner-turned samples and the frame table (about the raw data of the subint)
  long long ct = 0;
  for (const DevStation& st : p.st)
    ct += (long long)(d->blockspersend + 2) * st.L * st.nbits * st.nchan / 8 + st.framebytes * 4;
  return spec + v + blocks + res + tables + z + pcal + vpart + ct + ct / 8;
}

}  // namespace fxg
